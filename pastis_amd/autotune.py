"""Autotuner of the launch shapes (pa_tuning): times a sample batch of the
caller's own pairs under each candidate and keeps the fastest.

The candidates are the persistent DP workgroups per CU, the X-drop
workgroups per CU and the X-drop LDS ring size; results never depend on
them (checked by the GPU tests), only the speed does.
"""
import itertools
import json
import os
import time

import numpy as np

from ._lib import PAIR_DTYPE, RESULT_DTYPE, SCORE_DTYPE

DP_WG = (0, 16, 8)        # 0 = occupancy limit
XD_WG = (0, 8, 4)
XD_RING = (16, 32)


def candidates(mode_xdrop=True):
    """16-lanes-per-item X-drop with each DP grid and overflow ring size;
    one-lane-per-item X-drop with its grid and ring choices."""
    if not mode_xdrop:
        for dp in DP_WG:
            yield dict(dp_wg_per_cu=dp, xd_wg_per_cu=0, xd_ring=16,
                       xd_perlane=0)
        return
    for dp, ring in itertools.product(DP_WG, XD_RING):
        yield dict(dp_wg_per_cu=dp, xd_wg_per_cu=0, xd_ring=ring,
                   xd_perlane=0)
    for xd, ring in itertools.product(XD_WG, XD_RING):
        yield dict(dp_wg_per_cu=0, xd_wg_per_cu=xd, xd_ring=ring,
                   xd_perlane=1)


def tune(al, pairs, xdrop=True, sample=200000, repeats=2, seed=0,
         cache=None, log=None):
    """Picks the fastest tuning for `al` on a sample of `pairs`, installs it
    and returns (best, table).  cache: optional JSON file keyed by the
    sample's size and mode."""
    key = "%s:%d" % ("xdrop" if xdrop else "sw", min(sample, len(pairs)))
    if cache and os.path.exists(cache):
        with open(cache) as fh:
            known = json.load(fh)
        if key in known:
            al.set_tuning(**known[key])
            return known[key], []
    rng = np.random.default_rng(seed)
    pick = pairs if len(pairs) <= sample else \
        pairs[np.sort(rng.choice(len(pairs), sample, replace=False))]
    pick = np.ascontiguousarray(pick, PAIR_DTYPE)
    n = len(pick)
    out_t = RESULT_DTYPE if xdrop else SCORE_DTYPE
    dp = al.device_alloc(pick.nbytes)
    do = al.device_alloc(n * out_t.itemsize)
    table = []
    try:
        al.h2d(dp, pick)
        run = al.align_dev if xdrop else al.sw_scores_dev
        for cand in candidates(xdrop):
            al.set_tuning(**cand)
            run(dp, n, do)          # warm-up (first launch, allocations)
            al.synchronize()
            best = np.inf
            for _ in range(repeats):
                t = time.perf_counter()
                run(dp, n, do)
                al.synchronize()
                best = min(best, time.perf_counter() - t)
            table.append(dict(cand, ms=round(best * 1e3, 3)))
            if log:
                log("autotune %s: %.3f ms" % (cand, best * 1e3))
    finally:
        al.device_free(dp)
        al.device_free(do)
    best = min(table, key=lambda r: r["ms"])
    choice = {k: best[k] for k in ("dp_wg_per_cu", "xd_wg_per_cu", "xd_ring",
                                   "xd_perlane")}
    al.set_tuning(**choice)
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as fh:
                known = json.load(fh)
        known[key] = choice
        with open(cache, "w") as fh:
            json.dump(known, fh, indent=1)
    return choice, table
