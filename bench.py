"""Benchmark of the batched pairwise aligner on one MI355X.

A step is one pass of the aligner operator (PWAlign::aln_batch replacement,
pa_align_dev) over one batch of synthetic candidate pairs whose sequences and
pairs are already resident in HBM.  The default workload is the seed-and-
extend X-drop configuration: 20,000 proteins with log-normal lengths (median
250, sigma 0.6, clipped to 30..3000), 5,000,000 candidate pairs with up to two
shared 6-mer seeds, 30 % planted homologs at 30..95 % identity, BLOSUM62,
X = 49.  Each step runs the X-drop extension of both seeds, the global
alignment of the extended infixes with traceback statistics, the best-seed
choice and the coverage/identity predicate.

Prints one JSON line.  Besides the headline it carries, under "configs",
the other configurations measured the same way (fewer steps, each with its
dominant kernel's roofline and a CPU sample):
  C1  plain local Smith-Waterman: 2,000 proteins U[50,500], 100,000 pairs,
      score and end cell;
  C3  batched uneven-length Smith-Waterman: 200,000 proteins (log-normal
      median 300 plus a 1 % tail of 3,000-35,000 residues), 5e7 Zipf pairs
      resident in HBM, replayed as hipGraph-captured batches of 1e7 pairs;
  C5  the end-to-end candidate stream: 1,000,000 proteins (log-normal median
      280, 5 % in families of 2-50), k = 6 k-mer candidates (strict upper
      triangle, ckthr = 1) generated on the GPU, then learned filter + exact
      X-drop + triples against exact X-drop + triples on every candidate;
      the filter is trained on labels of a uniform sample of the stream.
--configs none skips them; --workload sw / sw3 runs C1 / C3 as the
headline instead.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("GCUPS + aligned pairs/s on 1 MI355X; learned-filter vs plain at "
          ">=99.5% recall")
# packed int16 VALU peak: 256 CU x 64 lanes x 2 halves x 2.4 GHz
PEAK_I16_TOPS = 78.6
PEAK_I32_TOPS = 39.3
OPS_PER_CELL = {"sw": 10, "swt": 10, "nwt": 9}   # algorithmic int ops / cell


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        n = min(n, int(env))
    return max(1, n)


FILTER_FLOP_PER_PAIR = 2 * (16 * 64 + 64 * 64 + 64)


# rocprof FETCH_SIZE / WRITE_SIZE summaries (KiB per dispatch) of a
# workload, with their copy-kernel calibration factors
TRAFFIC_PROFILES = {
    ("xdrop-lognormal", "as_called", 1.0):
        ("profiles/xdrop_5M_pairs_hbm_traffic.md", 2.0, 1.0),
}


def rocprof_name(kernel):
    """k_dp<nwt.c3.l64r8.i16> -> void pa::k_dp<8, 64, false, true, true>"""
    m = re.match(r"k_dp<(sw|swt|nwt)\.c\d\.l(\d+)r(\d+)\.(i16|i32)>", kernel)
    if not m:
        return None
    local, trace = {"sw": ("true", "false"), "swt": ("true", "true"),
                    "nwt": ("false", "true")}[m.group(1)]
    return "void pa::k_dp<%s, %s, %s, %s, %s>" % (
        m.group(3), m.group(2), local, trace,
        "true" if m.group(4) == "i16" else "false")


def profiled_traffic(kernel, key):
    """HBM bytes per launch of `kernel` from the committed rocprof summary
    of this workload (FETCH_SIZE and WRITE_SIZE tables, calibrated), or
    None when the workload or kernel was not profiled."""
    prof = TRAFFIC_PROFILES.get(key)
    name = rocprof_name(kernel)
    if not prof or not name:
        return None, None
    path, ff, fw = prof
    full = os.path.join(os.path.dirname(os.path.abspath(__file__)), path)
    if not os.path.exists(full):
        return None, None
    vals = {}
    for line in open(full):
        if line.startswith(name):
            parts = line.split()
            vals.setdefault("fetch" if "fetch" not in vals else "write",
                            float(parts[-1]))
    if len(vals) != 2:
        return None, None
    return int((vals["fetch"] * ff + vals["write"] * fw) * 1024), path


def roofline(kstats, st, step_ms_total, tuning=None, traffic_key=None):
    """Dominant kernel by device time: algorithmic ops / launch time.

    Candidates are the DP kernel classes (k_dp instantiations) and the X-drop
    extension kernel; ops per cell follow the fixed table (Gotoh local 10,
    global on the infix 9, X-drop 7 per evaluated cell).  traffic: HBM bytes
    per launch from the committed rocprof summary of the workload
    (traffic_key = (workload, gap model, scale)), when there is one."""
    cands = []
    for k in kstats:
        mode = k["name"].split(".")[0]
        cands.append(dict(kernel="k_dp<%s>" % k["name"], ms=k["ms"],
                          launches=k["launches"], cells=k["cells"],
                          ops=OPS_PER_CELL[mode],
                          peak=PEAK_I16_TOPS if k["name"].endswith("i16")
                          else PEAK_I32_TOPS))
    if st.get("ms_xdrop", 0) > 0:
        # the X-drop stage is timed as one: its kernels chained by
        # overflow lists, the last ones taking only the items that outgrow
        # the earlier ones' windows
        tn = tuning or {}
        chain = ["k_xdrop_coop"] if tn.get("xd_coop") else []
        chain.append("k_xdrop_pl<%d>" % tn.get("xd_ring", 16))
        chain += {0: [], 1: ["k_xdrop_band<8>", "k_xdrop_band<16>"],
                  2: ["k_xdrop_band<16>"]}[tn.get("xd_wave", 1)]
        chain.append("k_xdrop_pl<0>")
        xk = "X-drop stage: " + " -> ".join(chain)
        cands.append(dict(kernel=xk, ms=st["ms_xdrop"],
                          launches=None, cells=st["xdrop_cells"], ops=7,
                          peak=PEAK_I32_TOPS))
    if not cands:
        return None
    top = max(cands, key=lambda k: k["ms"])
    tops = top["ops"] * top["cells"] / (top["ms"] * 1e-3) / 1e12
    traffic, traffic_src = profiled_traffic(top["kernel"], traffic_key)
    return {"bound": "valu", "kernel": top["kernel"],
            "achieved": round(tops, 3), "peak": top["peak"], "unit": "Tops/s",
            "frac": round(tops / top["peak"], 4), "traffic": traffic,
            "traffic_source": traffic_src,
            "ops_per_cell": top["ops"], "cells": int(top["cells"]),
            "kernel_ms": round(top["ms"], 3),
            "launches": top["launches"],
            "share_of_timed_region": round(top["ms"] / step_ms_total, 3)}


def cpu_baseline_xdrop(w, pairs, p, budget_s, threads):
    from oracle import oracle as orc
    rng = np.random.default_rng(1234)
    # size the sample from a short probe so that the run takes ~budget_s
    probe = pairs[rng.choice(len(pairs), min(2000, len(pairs)),
                             replace=False)]
    t = time.time()
    orc.sxa_batch(w.codes, w.offsets, probe["r"], probe["c"], probe["seed"],
                  p, nthreads=threads)
    dt = max(time.time() - t, 1e-3)
    n = int(min(len(pairs), max(len(probe), len(probe) * budget_s / dt)))
    idx = np.sort(rng.choice(len(pairs), n, replace=False))
    sample = pairs[idx]
    t = time.time()
    ref = orc.sxa_batch(w.codes, w.offsets, sample["r"], sample["c"],
                        sample["seed"], p, nthreads=threads)
    dt = time.time() - t
    L = w.lengths
    nw_cells = 0
    for s in range(p.seed_cnt):
        e = ref["ext"][:, s]
        nw_cells += int(np.sum((e["end_h"] - e["beg_h"]).astype(np.int64) *
                               (e["end_v"] - e["beg_v"]).astype(np.int64)))
    x_cells = int(ref["ext"]["cells"].sum())
    cells = nw_cells + x_cells
    return idx, ref, {
        "value": round(cells / dt / 1e9, 4), "unit": "GCUPS",
        "pairs_per_s": round(n / dt, 1), "cores": threads, "kind": "port",
        "sample": "%d of the %d pairs drawn uniformly (seed 1234), %.1f s, "
                  "%d DP cells; oracle/pastis_oracle.c sxa path" %
                  (n, len(pairs), dt, cells)}


def cpu_baseline_sw(w, pairs, go, ge, budget_s, threads):
    from oracle import oracle as orc
    rng = np.random.default_rng(1234)
    probe = pairs[rng.choice(len(pairs), min(500, len(pairs)), replace=False)]
    t = time.time()
    orc.sw_batch(w.codes, w.offsets, probe["r"], probe["c"], go, ge,
                 nthreads=threads)
    dt = max(time.time() - t, 1e-3)
    n = int(min(len(pairs), max(500, len(probe) * budget_s / dt)))
    idx = np.sort(rng.choice(len(pairs), n, replace=False))
    sample = pairs[idx]
    t = time.time()
    ref = orc.sw_batch(w.codes, w.offsets, sample["r"], sample["c"], go, ge,
                       nthreads=threads)
    dt = time.time() - t
    L = w.lengths
    cells = int(np.sum(L[sample["r"]] * L[sample["c"]]))
    return idx, ref, {
        "value": round(cells / dt / 1e9, 4), "unit": "GCUPS",
        "pairs_per_s": round(n / dt, 1), "cores": threads, "kind": "port",
        "sample": "%d of the %d pairs drawn uniformly (seed 1234), %.1f s, "
                  "%d cells; oracle/pastis_oracle.c local score path" %
                  (n, len(pairs), dt, cells)}


def learned_filter(args, cfg, al, d_pairs, n, plain_out, plain_ms):
    """Train the prefilter on GPU-aligner labels of a separate workload of
    the same generator (X-drop: config 2 with 2M pairs at scale 1, seed 202;
    SW: config 1 at twice the scale, seed 201), calibrate tau for 99.5 %
    recall on a held-out tenth with a one-sided 95 % margin for the sampling
    error of that tenth (z = 1.645), then run filter + exact aligner on the
    benchmark pairs and compare
    with the plain operator (pa_align on every pair)."""
    from pastis_amd import prefilter, synth
    from pastis_amd._lib import PAIR_DTYPE, RESULT_DTYPE
    from pastis_amd.aligner import Aligner
    t = time.time()
    if args.workload == "xdrop":
        tw = synth.config2(scale=0.4 * args.scale, seed=202,
                           threads=host_cores())
    else:
        tw = synth.config1(scale=2 * args.scale, seed=201,
                           threads=host_cores())
    tal = Aligner(cfg)
    tal.load_seqs(tw.codes, tw.offsets)
    y = prefilter.labels_of(tal.align(tw.pairs))
    flt = prefilter.LearnedFilter(tal, seed=0)
    f = flt.features(tw.pairs)
    ntr = len(y) * 9 // 10
    t_train = time.time()
    losses = flt.train(f[:ntr], y[:ntr], epochs=4, batch=4096, lr=2e-3)
    t_train = time.time() - t_train
    tau = flt.calibrate(f[ntr:], y[ntr:], recall=0.995, z=1.645)
    params, mean, istd = tal.filter_get_model()
    tal.close()
    log("filter: trained on %d pairs (%d positive) in %.2f s, losses %s, "
        "tau %.4f (setup %.1f s)" % (ntr, int(y[:ntr].sum()), t_train,
                                     np.round(losses, 4).tolist(), tau,
                                     time.time() - t))
    # evaluation on the benchmark pairs against the exact labels
    al.filter_set_model(params, mean, istd)
    d_log = al.device_alloc(4 * n)
    al.filter_score_dev(d_pairs, n, d_log)
    z = np.zeros(n, np.float32)
    al.synchronize()
    al.d2h(z, d_log)
    # inference throughput of the fused features + MLP kernel alone
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        al.filter_score_dev(d_pairs, n, d_log)
    al.synchronize()
    score_ms = (time.perf_counter() - t0) * 1e3 / reps
    al.device_free(d_log)
    labels = prefilter.labels_of(plain_out)
    rep = prefilter.report(z, labels, tau)
    # timed: select + exact alignment of the kept pairs
    d_kept = al.device_alloc(n * PAIR_DTYPE.itemsize)
    d_idx = al.device_alloc(4 * n)
    d_res = al.device_alloc(n * RESULT_DTYPE.itemsize)

    def fstep():
        k = al.filter_select_dev(d_pairs, n, float(tau), d_kept, d_idx)
        al.align_dev(d_kept, k, d_res)
        return k

    for _ in range(args.warmup):
        fstep()
    al.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        k = fstep()
    al.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    res = np.zeros(k, RESULT_DTYPE)
    idx = np.zeros(k, np.uint32)
    al.d2h(res, d_res)
    al.d2h(idx, d_idx)
    exact = bool(np.array_equal(res, plain_out[idx]))
    for b in (d_kept, d_idx, d_res):
        al.device_free(b)
    return {
        "model": "MLP 16-64-64-1, bf16 weights/activations, fp32 accumulate",
        "train_pairs": int(ntr), "train_seconds": round(t_train, 3),
        "epochs": 4, "final_loss": round(float(losses[-1]), 5),
        "tau_logit": round(float(tau), 5),
        "recall": round(rep["recall"], 5),
        "precision": round(rep["precision"], 5),
        "kept_frac": round(rep["kept_frac"], 5),
        "band_frac": round(rep["band_frac"], 5),
        "ms_per_step": round(ms, 3),
        "pairs_per_s": round(n / (ms * 1e-3), 1),
        "speedup_vs_plain": round(plain_ms / ms, 3),
        "kept_results_identical_to_plain": exact,
        # 2 * (16*64 + 64*64 + 64) fp32 FLOP per pair (forward)
        "score_ms": round(score_ms, 3),
        "score_pairs_per_s": round(n / (score_ms * 1e-3), 1),
        "score_tflops": round(n * FILTER_FLOP_PER_PAIR / (score_ms * 1e-3)
                              / 1e12, 3),
    }


def roofline_of(al, steps_ms):
    st = al.stats()
    return roofline(al.kernel_stats(), st, steps_ms, al.tuning()), st


def measure_sw_config(args, which):
    """C1 (uniform lengths, one batch) or C3 (heavy tail, 5e7 pairs in
    hipGraph-captured batches of 1e7) as score-only local alignment with
    the pairs resident in HBM.  The roofline comes from one eager pass with
    kernel timers; the timed steps replay the graphs (C3) or run eagerly
    (C1)."""
    import ctypes as C
    from pastis_amd import synth
    from pastis_amd._lib import PAIR_DTYPE, SCORE_DTYPE
    from pastis_amd.aligner import Aligner, config_for_cli
    t = time.time()
    if which == "c1":
        w = synth.config1(scale=1.0, threads=host_cores())
        batch = len(w.pairs)
    else:
        w = synth.config3(scale=args.c3_scale, threads=host_cores())
        batch = 10_000_000
    pairs = w.pairs
    n = len(pairs)
    log("%s: %d sequences, %d pairs, %.3g cells, generated in %.1f s" %
        (w.name, len(w.offsets) - 1, n, float(w.cells()), time.time() - t))
    al = Aligner(config_for_cli(-11, -2, 6, 0, 2, args.gap_model))
    al.load_seqs(w.codes, w.offsets)
    d_pairs = al.device_alloc(pairs.nbytes)
    d_out = al.device_alloc(n * SCORE_DTYPE.itemsize)
    al.h2d(d_pairs, pairs)
    offs = list(range(0, n, batch))

    def run(graph):
        for off in offs:
            m = min(batch, n - off)
            dp = C.c_void_p(d_pairs.value + off * PAIR_DTYPE.itemsize)
            do = C.c_void_p(d_out.value + off * SCORE_DTYPE.itemsize)
            if graph:
                al.sw_scores_graph(dp, m, do)
            else:
                al.sw_scores_dev(dp, m, do)

    # eager passes with kernel timers: the default launch shapes, then the
    # autotuned ones (whose pass gives the roofline of the dominant kernel)
    run(False)
    al.synchronize()
    t0 = time.perf_counter()
    run(False)
    al.synchronize()
    default_ms = (time.perf_counter() - t0) * 1e3
    if not args.no_autotune:
        from pastis_amd import autotune
        # evenly spaced samples of the whole set (the heavy tail of C3 is
        # not spread evenly over the batches)
        autotune.tune(al, (d_pairs, n), xdrop=False, force=True, log=log)
        run(False)
        al.synchronize()
    al.reset_stats()
    t0 = time.perf_counter()
    run(False)
    al.synchronize()
    eager_ms = (time.perf_counter() - t0) * 1e3
    rl, st = roofline_of(al, eager_ms)
    reruns = int(st["rerun_pairs"])
    use_graph = which == "c3"
    run(use_graph)           # capture (C3)
    al.synchronize()
    steps = max(1, args.config_steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        run(use_graph)
    al.synchronize()
    el = time.perf_counter() - t0
    out = np.zeros(n, SCORE_DTYPE)
    al.d2h(out, d_out)
    cells = w.cells()
    res = {
        "workload": w.name, "sequences": len(w.offsets) - 1, "pairs": n,
        "batch_pairs": batch, "graph_batches": use_graph,
        "steps": steps, "ms_per_step": round(el * 1e3 / steps, 3),
        "value": round(cells * steps / el / 1e9, 3), "unit": "GCUPS",
        "pairs_per_s": round(n * steps / el, 1), "cells_per_step": cells,
        "eager_ms_per_step": round(eager_ms, 3),
        "eager_ms_per_step_default_tuning": round(default_ms, 3),
        "tuning": al.tuning(),
        "int32_rerun_pairs": reruns,
        "int32_rerun_rate": round(reruns / n, 8),
        "dtype": "int16 (packed; int32 re-run on saturation)",
        "roofline": rl,
    }
    al.device_free(d_pairs)
    al.device_free(d_out)
    al.close()
    if not args.no_cpu:
        go, ge = (-2, -11) if args.gap_model == "as_called" else (-11, -2)
        idx, ref, cb = cpu_baseline_sw(w, pairs, go, ge, args.cpu_seconds,
                                       host_cores())
        ok = (np.array_equal(out["score"][idx], ref[0]) and
              np.array_equal(out["r_end"][idx].astype(np.int32), ref[1]) and
              np.array_equal(out["c_end"][idx].astype(np.int32), ref[2]))
        res["cpu_baseline"] = cb
        res["parity_on_cpu_sample"] = "bit-exact" if ok else "MISMATCH"
    return res


def c5_train_filter(al, sample):
    """Learned filter for the candidate stream, trained on GPU-aligner labels
    of a uniform sample of the stream's own candidates; tau calibrated on a
    held-out 30 % for 99.5 % recall with a one-sided 95 % margin.  Installs
    the model on `al` and returns (tau, info)."""
    from pastis_amd import prefilter
    t = time.time()
    y = prefilter.labels_of(al.align(sample))
    flt = prefilter.LearnedFilter(al, seed=0)
    f = flt.features(sample)
    rng = np.random.default_rng(7)
    perm = rng.permutation(len(sample))
    f, y = f[perm], y[perm]
    ntr = len(y) * 7 // 10
    t1 = time.time()
    losses = flt.train(f[:ntr], y[:ntr], epochs=4, batch=4096, lr=2e-3)
    t_train = time.time() - t1
    tau = flt.calibrate(f[ntr:], y[ntr:], recall=0.995, z=1.645)
    info = {"train_pairs": int(ntr), "train_positives": int(y[:ntr].sum()),
            "calibration_pairs": int(len(y) - ntr),
            "train_seconds": round(t_train, 3),
            "final_loss": round(float(losses[-1]), 5),
            "tau_logit": round(float(tau), 5),
            "setup_seconds": round(time.time() - t, 1)}
    log("C5 filter: %s" % info)
    return tau, info


def measure_c5(args):
    """End-to-end candidate stream on one GPU (see the module docstring)."""
    import ctypes as C
    from pastis_amd import io, prefilter, synth
    from pastis_amd._lib import PAIR_DTYPE, RESULT_DTYPE, TRIPLE_DTYPE
    from pastis_amd.aligner import Aligner, config_for_cli
    cfg = config_for_cli(-11, -2, 6, 49, 2, args.gap_model)
    t = time.time()
    w = synth.config5(scale=args.c5_scale)
    nseq = len(w.offsets) - 1
    log("C5: %d sequences (%d residues) generated in %.1f s" %
        (nseq, len(w.codes), time.time() - t))
    al = Aligner(cfg)
    al.load_seqs(w.codes, w.offsets)
    # candidate generation: column blocks, kept resident
    kmer = dict(k=6, symmetric="upper", ckthr=1)
    nblk = 4
    bounds = [(nseq * i // nblk, nseq * (i + 1) // nblk) for i in range(nblk)]
    al.synchronize()
    t0 = time.perf_counter()
    counts = [al.kmer_candidates_dev(None, 0, c_begin=a, c_end=b, **kmer)
              for a, b in bounds]
    gen_count_s = time.perf_counter() - t0
    ncand = int(sum(counts))
    d_all = al.device_alloc(max(ncand, 1) * PAIR_DTYPE.itemsize)
    off = 0
    t0 = time.perf_counter()
    for (a, b), m in zip(bounds, counts):
        dp = C.c_void_p(d_all.value + off * PAIR_DTYPE.itemsize)
        al.kmer_candidates_dev(dp, m, c_begin=a, c_end=b, **kmer)
        off += m
    gen_s = gen_count_s + (time.perf_counter() - t0)
    log("C5: %d candidates in %.2f s" % (ncand, gen_s))
    # the filter learns from a uniform sample of the stream itself (at most
    # 4,000,000 candidates, aligned exactly for their labels)
    full = np.zeros(ncand, PAIR_DTYPE)
    if ncand:
        al.d2h(full, d_all)
    rng = np.random.default_rng(11)
    sidx = np.sort(rng.choice(ncand, min(ncand, 4_000_000), replace=False))
    tau, finfo = c5_train_filter(al, full[sidx])
    batch = 10_000_000
    b = min(batch, max(ncand, 1))
    d_res = al.device_alloc(b * RESULT_DTYPE.itemsize)
    d_kept = al.device_alloc(b * PAIR_DTYPE.itemsize)
    d_idx = al.device_alloc(b * 4)
    d_tri = al.device_alloc(max(ncand, 1) * TRIPLE_DTYPE.itemsize)
    flags = np.zeros(ncand, np.uint16)
    kept_pass = [0]

    def stream(filtered, collect=False):
        ntri = 0
        for o in range(0, ncand, batch):
            m = min(batch, ncand - o)
            dp = C.c_void_p(d_all.value + o * PAIR_DTYPE.itemsize)
            if filtered:
                m = al.filter_select_dev(dp, m, float(tau), d_kept, d_idx)
                dp = d_kept
            if m == 0:
                continue
            al.align_dev(dp, m, d_res)
            dt = C.c_void_p(d_tri.value + ntri * TRIPLE_DTYPE.itemsize)
            ntri += io.triples_dev(al, dp, d_res, m, dt)
            if collect:
                r = np.zeros(m, RESULT_DTYPE)
                al.d2h(r, d_res)
                if filtered:
                    kept_pass[0] += int(np.sum(prefilter.labels_of(r)))
                else:
                    flags[o:o + m] = r["flags"]
        return ntri

    # untimed passes: labels of every candidate, the filter's kept positives
    ntri_plain = stream(False, collect=True)
    ntri_filt = stream(True, collect=True)
    npos = int(np.sum((flags & 1) != 0))
    al.synchronize()
    steps = max(1, args.config_steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        stream(False)
    al.synchronize()
    plain_s = (time.perf_counter() - t0) / steps
    t0 = time.perf_counter()
    for _ in range(steps):
        stream(True)
    al.synchronize()
    filt_s = (time.perf_counter() - t0) / steps
    tri = np.zeros(ntri_filt, TRIPLE_DTYPE)
    if ntri_filt:
        al.d2h(tri, d_tri)
    res = {
        "workload": w.name, "sequences": nseq, "residues": int(len(w.codes)),
        "kmer": "k=6, strict upper triangle, ckthr=1, pdefault alphabet",
        "candidates": ncand, "batch_pairs": batch,
        "candidate_generation_s": round(gen_s, 3),
        "candidate_generation_pairs_per_s": round(ncand / max(gen_s, 1e-9), 1),
        "passing_pairs": npos, "triples_plain": ntri_plain,
        "triples_filtered": ntri_filt,
        "plain": {"s_per_step": round(plain_s, 4),
                  "pairs_per_s": round(ncand / plain_s, 1)},
        "filter_exact": {"s_per_step": round(filt_s, 4),
                         "pairs_per_s": round(ncand / filt_s, 1)},
        "speedup_filter_vs_plain": round(plain_s / filt_s, 3),
        "recall": round(kept_pass[0] / npos, 5) if npos else 1.0,
        "end_to_end_pairs_per_s_plain": round(ncand / (gen_s + plain_s), 1),
        "end_to_end_pairs_per_s_filter":
            round(ncand / (gen_s + filt_s), 1),
        "filter": finfo,
    }
    pairs = None
    if not args.no_cpu and ncand:
        # the CPU sample is drawn from the candidates
        rng = np.random.default_rng(13)
        pairs = full[np.sort(rng.choice(ncand, min(ncand, 2_000_000),
                                        replace=False))]
    del full
    for x in (d_all, d_res, d_kept, d_idx, d_tri):
        al.device_free(x)
    al.close()
    if pairs is not None:
        from oracle import oracle as orc
        g = ({"gap_open": -2, "gap_ext": -11, "xdrop_gap": -11}
             if args.gap_model == "as_called" else
             {"gap_open": -11, "gap_ext": -2, "xdrop_gap": -2})
        p = orc.params(g["gap_open"], g["gap_ext"], g["xdrop_gap"], 49, 6, 2,
                       0.7, 30)
        wk = synth.Workload(w.name, w.codes, w.offsets, pairs, w.family, {})
        _, ref, cb = cpu_baseline_xdrop(wk, pairs, p, args.cpu_seconds,
                                        host_cores())
        cb["unit"] = "candidate pairs/s (X-drop + statistics)"
        cb["value"] = cb.pop("pairs_per_s")
        res["cpu_baseline"] = cb
    log("C5: %s" % {k: v for k, v in res.items() if k != "filter"})
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["xdrop", "sw", "sw3"],
                    default="xdrop")
    ap.add_argument("--graph", action="store_true",
                    help="sw / sw3: replay each step from a captured HIP graph")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--gap-model", default="as_called",
                    choices=["as_called", "as_documented"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", type=float, default=0.0, metavar="GIB",
                    help="only run 3 device-to-device copies of GIB GiB "
                         "(HBM counter calibration) and exit")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes")
    ap.add_argument("--set-tuning", default="",
                    help="launch shapes to install, e.g. tb_stage=1,"
                         "dp_shape=0:1:2:3 (profiling runs)")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip timing pa_align_graph replays")
    ap.add_argument("--no-operator", action="store_true",
                    help="skip timing HipAligner::aln_batch on host tuples")
    ap.add_argument("--no-filter", action="store_true",
                    help="skip the learned-filter vs plain comparison")
    ap.add_argument("--configs", default="c1,c3,c5",
                    help="extra configurations measured after the headline "
                         "(comma list of c1, c3, c5, or none)")
    ap.add_argument("--c3-scale", type=float, default=1.0)
    ap.add_argument("--c5-scale", type=float, default=1.0)
    ap.add_argument("--config-steps", type=int, default=2,
                    help="timed steps of each extra configuration")
    args = ap.parse_args()
    if args.copy_probe > 0:
        import torch
        nb = int(args.copy_probe * (1 << 30))
        a = torch.empty(nb, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied %d bytes 3 times" % nb)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU "
              "path (pairs are independent and would shard by similarity-matrix tile, one GPU per tile)", file=sys.stderr)
        sys.exit(2)

    from pastis_amd import synth
    from pastis_amd._lib import PAIR_DTYPE, RESULT_DTYPE, SCORE_DTYPE
    from pastis_amd.aligner import Aligner, config_for_cli

    t = time.time()
    if args.workload == "xdrop":
        w = synth.config2(scale=args.scale, threads=host_cores())
        cfg = config_for_cli(-11, -2, 6, 49, 2, args.gap_model)
    elif args.workload == "sw":
        w = synth.config1(scale=args.scale, threads=host_cores())
        cfg = config_for_cli(-11, -2, 6, 0, 2, args.gap_model)
    else:
        w = synth.config3(scale=args.scale, threads=host_cores())
        cfg = config_for_cli(-11, -2, 6, 0, 2, args.gap_model)
    pairs = w.pairs
    n = len(pairs)
    log("workload %s: %d sequences, %d pairs, generated in %.1f s" %
        (w.name, len(w.offsets) - 1, n, time.time() - t))

    al = Aligner(cfg)
    al.load_seqs(w.codes, w.offsets)
    d_pairs = al.device_alloc(n * PAIR_DTYPE.itemsize)
    out_dtype = RESULT_DTYPE if args.workload == "xdrop" else SCORE_DTYPE
    d_out = al.device_alloc(n * out_dtype.itemsize)
    al.h2d(d_pairs, pairs)
    def step():
        if args.workload == "xdrop":
            al.align_dev(d_pairs, n, d_out)
        elif args.graph:
            al.sw_scores_graph(d_pairs, n, d_out)
        else:
            al.sw_scores_dev(d_pairs, n, d_out)

    # one step with the default launch shapes (after a warm-up step), for
    # the autotuner's gain
    step()
    al.synchronize()
    t0 = time.perf_counter()
    step()
    al.synchronize()
    default_ms = (time.perf_counter() - t0) * 1e3
    tune_s = None
    if not args.no_autotune:
        from pastis_amd import autotune
        _, tune_s = autotune.tune(al, (d_pairs, n),
                                  xdrop=args.workload == "xdrop", log=log)
    if args.set_tuning:
        # explicit launch shapes (after any autotuning): k=v,k=v with
        # dp_edge / dp_shape given as a:b:c
        kw = {}
        for item in args.set_tuning.split(","):
            k, v = item.split("=")
            kw[k] = [int(x) for x in v.split(":")] if ":" in v else int(v)
        al.set_tuning(**kw)
    tuning = al.tuning()

    for i in range(args.warmup):
        step()
        al.synchronize()
        log("warmup step %d done" % i)
    al.reset_stats()
    al.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
    al.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    st = al.stats()
    ks = al.kernel_stats()
    log("timed %d steps in %.3f s" % (args.steps, elapsed))
    for k in sorted(ks, key=lambda k: -k["ms"]):
        log("  %-16s launches %5d  ms %9.2f  cells %.3e" %
            (k["name"], k["launches"], k["ms"], k["cells"]))
    cells = st["dp_cells"] + st["xdrop_cells"]
    gcups = cells / elapsed / 1e9
    ms_per_step = elapsed * 1e3 / args.steps
    out = np.zeros(n, out_dtype)
    al.d2h(out, d_out)
    graph = None
    if args.workload == "xdrop" and not args.no_graph:
        # the same steps replayed from a captured HIP graph (pa_align_graph:
        # trace store chunked on the device, no host synchronisation)
        d_gout = al.device_alloc(n * out_dtype.itemsize)
        al.align_graph(d_pairs, n, d_gout)     # eager run + capture
        al.synchronize()
        al.align_graph(d_pairs, n, d_gout)     # first replay
        al.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            al.align_graph(d_pairs, n, d_gout)
        al.synchronize()
        g_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        gout = np.zeros(n, out_dtype)
        al.d2h(gout, d_gout)
        al.device_free(d_gout)
        graph = {"ms_per_step": round(g_ms, 3),
                 "results_identical_to_eager": bool(np.array_equal(gout, out))}
        log("graph replays: %.1f ms per step" % g_ms)

    line = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "int16 (packed; int32 re-run on saturation)",
        "data": "synthetic",
        "config": {"workload": w.name, "sequences": len(w.offsets) - 1,
                   "pairs": n, "gap_model": args.gap_model,
                   "scale": args.scale,
                   "xdrop": 49 if args.workload == "xdrop" else None,
                   "seed_count": 2, "k": 6, "matrix": "BLOSUM62",
                   "tuning": tuning},
        "ms_per_step_default_tuning": round(default_ms, 3),
        "graph_replay": graph,
        "autotune_seconds": None if tune_s is None else round(tune_s, 2),
        "pairs_per_s": round(n * args.steps / elapsed, 1),
        "cells_per_step": int(cells / args.steps),
        "xdrop_cells_per_step": int(st["xdrop_cells"] / args.steps),
        "passing_pairs": int(((out["flags"] & 1) != 0).sum())
        if args.workload == "xdrop" else None,
        "roofline": roofline(ks, st, elapsed * 1e3, tuning,
                             (w.name, args.gap_model, args.scale)),
        "device_ms_per_step": {k: round(st[k] / args.steps, 3) for k in
                               ("ms_sw", "ms_xdrop", "ms_nw", "ms_trace")},
        "cpu_baseline": None,
    }
    if args.workload == "xdrop" and not args.no_filter:
        line["learned_filter"] = learned_filter(args, cfg, al, d_pairs, n, out,
                                                ms_per_step)
    elif args.workload == "sw" and not args.no_filter:
        # the plain operator: local alignment with traceback statistics and
        # the coverage/identity predicate on every pair
        d_res = al.device_alloc(n * RESULT_DTYPE.itemsize)
        for _ in range(args.warmup):
            al.align_dev(d_pairs, n, d_res)
        al.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            al.align_dev(d_pairs, n, d_res)
        al.synchronize()
        op_ms = (time.perf_counter() - t0) * 1e3 / args.steps
        res = np.zeros(n, RESULT_DTYPE)
        al.d2h(res, d_res)
        al.device_free(d_res)
        lf = learned_filter(args, cfg, al, d_pairs, n, res, op_ms)
        lf["plain_operator_ms_per_step"] = round(op_ms, 3)
        line["learned_filter"] = lf
    if not args.no_cpu:
        threads = host_cores()
        if args.workload == "xdrop":
            from oracle import oracle as orc
            g = ({"gap_open": -2, "gap_ext": -11, "xdrop_gap": -11}
                 if args.gap_model == "as_called" else
                 {"gap_open": -11, "gap_ext": -2, "xdrop_gap": -2})
            p = orc.params(g["gap_open"], g["gap_ext"], g["xdrop_gap"], 49, 6,
                           2, 0.7, 30)
            idx, ref, cb = cpu_baseline_xdrop(w, pairs, p, args.cpu_seconds,
                                              threads)
            ok = (np.array_equal(out["score"][idx], ref["score"]) and
                  np.array_equal(out["score_aln"][idx], ref["score_aln"]) and
                  np.array_equal(out["cols"][idx].astype(np.int64),
                                 ref["cols"].astype(np.int64)))
        else:
            go, ge = (-2, -11) if args.gap_model == "as_called" else (-11, -2)
            idx, ref, cb = cpu_baseline_sw(w, pairs, go, ge, args.cpu_seconds,
                                           threads)
            ok = (np.array_equal(out["score"][idx], ref[0]) and
                  np.array_equal(out["r_end"][idx].astype(np.int32), ref[1]) and
                  np.array_equal(out["c_end"][idx].astype(np.int32), ref[2]))
        line["cpu_baseline"] = cb
        line["parity_on_cpu_sample"] = "bit-exact" if ok else "MISMATCH"
    al.device_free(d_pairs)
    al.device_free(d_out)
    al.close()
    if args.workload == "xdrop" and not args.no_operator:
        # the PWAlign operator on host tuples (pack, page-locked staging,
        # alignment, score_aln write-back), same pairs and launch shapes
        from pastis_amd import pwalign
        t = time.time()
        op_ms, op_passed = pwalign.bench_aln_batch(
            1, w.codes, w.offsets, pairs, gap_model=args.gap_model,
            autotune=not args.no_autotune, warmup=1,
            steps=max(1, min(args.steps, 5)))
        line["operator_path"] = {
            "entry": "HipAligner::aln_batch (CommonKmerLoc tuples in host "
                     "memory)",
            "ms_per_call": round(op_ms, 3),
            "pairs_per_s": round(n / (op_ms * 1e-3), 1),
            "passed": int(op_passed),
            "passed_matches_device_path": int(op_passed) ==
            line["passing_pairs"],
            "vs_device_resident_ms_per_step": round(op_ms / ms_per_step, 3)}
        log("operator path: %.1f ms per aln_batch call (%.1f s)" %
            (op_ms, time.time() - t))
    extra = [c.strip() for c in args.configs.split(",")
             if c.strip() and c.strip() != "none"]
    if extra:
        line["configs"] = {}
    for c in extra:
        t = time.time()
        try:
            if c == "c1":
                line["configs"]["C1"] = measure_sw_config(args, "c1")
            elif c == "c3":
                line["configs"]["C3"] = measure_sw_config(args, "c3")
            elif c == "c5":
                line["configs"]["C5"] = measure_c5(args)
            else:
                raise ValueError("unknown configuration " + c)
        except Exception as e:   # keep the headline line if one fails
            line["configs"][c.upper()] = {"error": repr(e)}
        log("config %s done in %.1f s" % (c, time.time() - t))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
