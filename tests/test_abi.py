"""The C-ABI library loads, exports every declared entry point, and its host
logic behaves (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pastis_amd import _lib
from pastis_amd.aligner import kmer_prune, config_for_cli

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "pastis_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pa_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    decl = declared_entry_points()
    assert len(decl) >= 19
    assert sorted(_lib.EXPORTS) == decl
    for name in decl:
        assert hasattr(L, name), name


def test_operator_layer_exports_its_entry_point():
    text = open(os.path.join(ROOT, "include", "pastis_pw.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decl = sorted(set(re.findall(r"\b(pw_[a-z0-9_]+)\s*\(", text)))
    assert decl == ["pw_bench_aln_batch", "pw_run_block"]
    from pastis_amd import pwalign
    L = pwalign.lib()
    for name in decl:
        assert hasattr(L, name), name


def test_strerror_names_every_status():
    L = _lib.lib()
    for code in range(11):
        s = L.pa_strerror(code).decode()
        assert s and s != "unknown status"
    assert L.pa_strerror(999).decode() == "unknown status"


def test_config_for_cli_gap_models():
    c = config_for_cli(-11, -2, 6, 49, 2, "as_called")
    # SeqAn's Score(gapExtend, gapOpen) receives (-11, -2) positionally
    assert (c.gap_open, c.gap_ext, c.xdrop_gap) == (-2, -11, -11)
    assert c.mode == _lib.PA_MODE_XDROP and c.xdrop == 49
    assert c.seed_len == 6 and c.seed_cnt == 2
    assert c.ani_thr == 30 and c.cov_thr == pytest.approx(0.7)
    d = config_for_cli(-11, -2, 6, 0, 2, "as_documented")
    assert (d.gap_open, d.gap_ext, d.xdrop_gap) == (-11, -2, -2)
    assert d.mode == _lib.PA_MODE_SW


def test_kmer_prune_matches_reference_predicate():
    rng = np.random.default_rng(3)
    p = np.zeros(1000, _lib.PAIR_DTYPE)
    p["count"] = rng.integers(1, 6, 1000)
    p["kscore"] = rng.integers(-1, 40, 1000)
    for ckthr, mosthr in [(0, -1.0), (1, -1.0), (2, 25.0), (0, 30.5)]:
        keep = kmer_prune(p, ckthr, mosthr)
        # DistPWRunner.cpp:48-54: prune iff count <= ckthr || score <= mosthr
        exp = ~((p["count"].astype(int) <= ckthr) |
                (p["kscore"].astype(np.float32) <= np.float32(mosthr)))
        assert (keep == exp).all()


def test_create_rejects_bad_config_without_touching_gpu():
    L = _lib.lib()
    cfg = config_for_cli(-11, -2, 6, 49, 2, "as_called")
    cfg.gap_open = 0   # zero penalties are rejected before any device call
    ctx = C.c_void_p()
    assert L.pa_create(C.byref(cfg), C.byref(ctx)) == 1
    cfg = config_for_cli(-11, -2, 6, 49, 3, "as_called")  # seed_cnt 3
    assert L.pa_create(C.byref(cfg), C.byref(ctx)) == 1
