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
                       xd_coop=0)
        return
    for dp, ring in itertools.product(DP_WG, XD_RING):
        yield dict(dp_wg_per_cu=dp, xd_wg_per_cu=0, xd_ring=ring, xd_coop=1)
    for xd, ring in itertools.product(XD_WG, XD_RING):
        yield dict(dp_wg_per_cu=0, xd_wg_per_cu=xd, xd_ring=ring, xd_coop=0)


def tune(al, pairs, xdrop=True, sample=1000000, repeats=2, seed=0,
         cache=None, log=None):
    """Picks the fastest tuning for `al` on a sample of `pairs`, installs it
    and returns (best, table).  The X-drop settings are chosen by the X-drop
    kernels' device time and the DP setting by the DP kernels' device time
    (each measured with the events of pa_get_stats), so that one stage's
    noise does not decide the other's.  cache: optional JSON file keyed by
    the sample's size and mode."""
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
            best = dict(ms=np.inf, ms_dp=np.inf, ms_xdrop=np.inf)
            for _ in range(repeats):
                al.reset_stats()
                t = time.perf_counter()
                run(dp, n, do)
                al.synchronize()
                el = time.perf_counter() - t
                st = al.stats()
                best["ms"] = min(best["ms"], el * 1e3)
                best["ms_dp"] = min(best["ms_dp"], st["ms_nw"] + st["ms_sw"])
                best["ms_xdrop"] = min(best["ms_xdrop"], st["ms_xdrop"])
            table.append(dict(cand, **{k: round(v, 3) for k, v in best.items()}))
            if log:
                log("autotune %s: %.3f ms (dp %.3f, xdrop %.3f)" %
                    (cand, best["ms"], best["ms_dp"], best["ms_xdrop"]))
    finally:
        al.device_free(dp)
        al.device_free(do)
    # the DP grid from the DP time, the X-drop shape from the X-drop time
    bdp = min(table, key=lambda r: r["ms_dp"])
    bxd = min(table, key=lambda r: r["ms_xdrop"]) if xdrop else bdp
    choice = {"dp_wg_per_cu": bdp["dp_wg_per_cu"],
              "xd_wg_per_cu": bxd["xd_wg_per_cu"], "xd_ring": bxd["xd_ring"],
              "xd_coop": bxd["xd_coop"]}
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
