"""The library's kernels and host code on the CPU, under AddressSanitizer and
UndefinedBehaviorSanitizer, against the oracle.

tests/emu compiles the HIP sources of libpastis_hip for the host, with every
wavefront emulated by 64 threads that meet at each cross-lane operation
(DPP, ballot, shuffle, readlane).  A cross-lane operation that only some
lanes of a wave reach (the pattern behind a GPU hang: a lane-0-only branch
ahead of a wave-wide step) leaves the other lanes waiting, so such a
regression shows up here as a time-out instead of on the GPU.  The driver
(tests/emu/driver.cpp) runs one batch through the C ABI: --sfa with
traceback statistics, and --sxa with both X-drop kernels (one lane per item,
and 16 lanes per item).  Results must equal the oracle's bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import oracle_params, small_config1, small_config2
from oracle import oracle as orc
from pastis_amd._lib import EXTENT_DTYPE, PAIR_DTYPE, RESULT_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "emu")
OUT = os.path.join(EMU, "build_san")
DRIVER = os.path.join(OUT, "driver")
SAN = "-fsanitize=address,undefined -fno-sanitize-recover=undefined"


@pytest.fixture(scope="module")
def driver():
    subprocess.run(["make", "-s", "-j8", "-C", EMU, "OUT=" + OUT,
                    "SAN=" + SAN, DRIVER], check=True, timeout=900)
    return DRIVER


def run(driver, tmp_path, w, pairs, mode, xdrop=0, coop=0,
        gap_model=0, shapes=None, xd_wave=None):
    inp = tmp_path / "in.bin"
    outp = tmp_path / "out.bin"
    hdr = np.array([mode, -11, -2, 6, xdrop, 2, gap_model, coop], np.int32)
    offs = np.ascontiguousarray(w.offsets, np.uint64)
    with open(inp, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([len(offs) - 1, len(pairs)], np.uint64).tobytes())
        f.write(offs.tobytes())
        f.write(np.ascontiguousarray(w.codes, np.uint8).tobytes())
        f.write(np.ascontiguousarray(pairs, PAIR_DTYPE).tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    if shapes is not None:
        env["EMU_SHAPES"] = ",".join(str(v) for v in shapes)
    if xd_wave is not None:
        env["EMU_XDWAVE"] = str(xd_wave)
    r = subprocess.run([driver, str(inp), str(outp)], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    raw = open(outp, "rb").read()
    n = len(pairs)
    res = np.frombuffer(raw[:n * RESULT_DTYPE.itemsize], RESULT_DTYPE)
    ext = cells = None
    if mode == 1:
        o = n * RESULT_DTYPE.itemsize
        ext = [np.frombuffer(raw[o + s * n * 8:o + (s + 1) * n * 8],
                             EXTENT_DTYPE) for s in range(2)]
        cells = int(np.frombuffer(raw[o + 2 * n * 8:o + 2 * n * 8 + 8],
                                  np.uint64)[0])
    return res, ext, cells


# Non-default launch shapes: 64-lane kernels on short rows and 16-lane
# multi-pass kernels (128 rows per pass) on long ones, class edges that send
# most tasks through several passes, LDS-staged tracebacks, each class's
# walks beside the next classes' DP; then the 32- and 8-lane kernels (eight
# and two tasks per wave) with multi-pass tasks and the walks beside the last
# class's DP.
SHAPES = [None, (3, 2, 1, 0, 16, 40, 100, 1, 1),
          (4, 5, 4, 5, 16, 40, 100, 1, 2)]


@pytest.mark.parametrize("shapes", SHAPES)
def test_emulated_sfa_under_sanitizers(driver, tmp_path, shapes):
    w = small_config1(n_seq=20, n_pairs=20, seed=71)
    got, _, _ = run(driver, tmp_path, w, w.pairs, 0, shapes=shapes)
    ref = orc.sfa_batch(w.codes, w.offsets, w.pairs["r"], w.pairs["c"],
                        oracle_params("as_called"))
    for gf, rf in [("score", "score"), ("r_end", "v_end"), ("c_end", "h_end"),
                   ("r_beg", "v_beg"), ("c_beg", "h_beg"),
                   ("matches", "matches"), ("cols", "cols"),
                   ("gap_opens", "gap_opens")]:
        assert np.array_equal(got[gf].astype(np.int64),
                              ref[rf].astype(np.int64)), gf
    assert np.array_equal(got["score_aln"], ref["score_aln"])


# the as-documented gap reading (X-drop gap -2) grows bands of hundreds of
# columns: those items leave the per-lane kernel for the wave-per-item one
# (xd_wave 2: the 256-column band kernel alone)
@pytest.mark.parametrize("coop,shapes,model,xd_wave", [
    (0, None, "as_called", None), (1, None, "as_called", None),
    (0, SHAPES[1], "as_called", None), (0, SHAPES[2], "as_called", None),
    (0, None, "as_documented", None),
    (0, None, "as_documented", 2)])
def test_emulated_sxa_under_sanitizers(driver, tmp_path, coop, shapes, model,
                                       xd_wave):
    w = small_config2(n_seq=20, n_pairs=16, seed=72)
    # the wide bands of the as-documented reading run through the 16-lane
    # band kernel, whose emulation is slow: three pairs (12 work items)
    pairs = w.pairs if model == "as_called" else w.pairs[:3]
    got, ext, cells = run(driver, tmp_path, w, pairs, 1, xdrop=49,
                          coop=coop, shapes=shapes,
                          gap_model=0 if model == "as_called" else 1,
                          xd_wave=xd_wave)
    ref = orc.sxa_batch(w.codes, w.offsets, pairs["r"], pairs["c"],
                        pairs["seed"], oracle_params(model))
    for gf in ("score", "matches", "cols", "gap_opens"):
        assert np.array_equal(got[gf].astype(np.int64),
                              ref[gf].astype(np.int64)), gf
    assert np.array_equal(got["score_aln"], ref["score_aln"])
    for s in range(2):
        e = ref["ext"][:, s]
        assert np.array_equal(ext[s]["r_beg"], e["beg_v"])
        assert np.array_equal(ext[s]["r_end"], e["end_v"])
        assert np.array_equal(ext[s]["c_beg"], e["beg_h"])
        assert np.array_equal(ext[s]["c_end"], e["end_h"])
    # evaluated X-drop cells of the pa_align call
    assert cells == int(ref["ext"]["cells"].sum())


def test_emulated_xdrop_int16_cells(driver, tmp_path):
    """X = 200 is beyond the int8 cell encoding of the per-lane X-drop
    kernel (X <= 160): its int16 variant, and the band kernels for the items
    that outgrow 16 columns, against the oracle under the sanitizers."""
    w = small_config2(n_seq=20, n_pairs=16, seed=73)
    pairs = w.pairs[:2]   # eight work items: the band kernels emulate slowly
    got, ext, cells = run(driver, tmp_path, w, pairs, 1, xdrop=200)
    ref = orc.sxa_batch(w.codes, w.offsets, pairs["r"], pairs["c"],
                        pairs["seed"], oracle_params("as_called", xdrop=200))
    for gf in ("score", "matches", "cols", "gap_opens"):
        assert np.array_equal(got[gf].astype(np.int64),
                              ref[gf].astype(np.int64)), gf
    for s in range(2):
        e = ref["ext"][:, s]
        assert np.array_equal(ext[s]["r_beg"], e["beg_v"])
        assert np.array_equal(ext[s]["r_end"], e["end_v"])
        assert np.array_equal(ext[s]["c_beg"], e["beg_h"])
        assert np.array_equal(ext[s]["c_end"], e["end_h"])
    assert cells == int(ref["ext"]["cells"].sum())


def run_kmer(driver, tmp_path, strs, k=3, stride=1, alph=0, sym=0, ckthr=0):
    inp = tmp_path / "k_in.bin"
    outp = tmp_path / "k_out.bin"
    res = np.frombuffer("".join(strs).encode(), np.uint8)
    offs = np.zeros(len(strs) + 1, np.uint64)
    offs[1:] = np.cumsum([len(s) for s in strs])
    hdr = np.array([2, k, stride, alph, sym, ckthr, 1, 0], np.int32)
    with open(inp, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([len(strs), 0], np.uint64).tobytes())
        f.write(offs.tobytes())
        f.write(res.tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([driver, str(inp), str(outp)], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    raw = open(outp, "rb").read()
    n = int(np.frombuffer(raw[:8], np.uint64)[0])
    return np.frombuffer(raw[8:8 + n * PAIR_DTYPE.itemsize], PAIR_DTYPE), \
        res, offs


@pytest.mark.parametrize("alph,sym,ckthr", [(0, 0, 0), (1, 1, 1), (5, 2, 0)])
def test_emulated_kmer_candidates_under_sanitizers(driver, tmp_path, alph,
                                                   sym, ckthr):
    import random
    rng = random.Random(alph * 7 + sym)
    strs = ["".join(rng.choice("ACDEWKLu") for _ in range(rng.randint(0, 30)))
            for _ in range(14)]
    got, res, offs = run_kmer(driver, tmp_path, strs, k=3, alph=alph,
                              sym=sym, ckthr=ckthr)
    name = {v: k for k, v in orc.ALPHABETS.items()}[alph]
    symn = {v: k for k, v in orc.SYMMETRIC.items()}[sym]
    ref = orc.kmer_candidates(res, offs, k=3, alph=name, sym=symn)
    ref = ref[~(ref["count"] <= ckthr)]
    assert len(got) == len(ref)
    for f in ("r", "c", "count", "kscore", "seed"):
        assert np.array_equal(got[f], ref[f]), f
