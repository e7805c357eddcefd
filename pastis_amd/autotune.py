"""Autotuning of the launch shapes (pa_tuning) through the library's own
tuner, pa_autotune: it times a sample of the caller's pairs under candidate
row-class edges, per-class kernel shapes, grid sizes, traceback staging and
X-drop kernels, each chosen by the device time of the stage it affects, and
caches the choice per device, target and alignment mode.  Results never
depend on the tuning (checked by the GPU tests), only the speed does.
"""
import time

import numpy as np

from ._lib import PAIR_DTYPE


def tune(al, pairs, xdrop=True, sample=262144, force=False, log=None):
    """Tunes `al` on `pairs` (a host PAIR_DTYPE array or a (device pointer,
    n) tuple); returns (choice, seconds)."""
    t = time.time()
    if isinstance(pairs, tuple):
        dp, n = pairs
        own = False
    else:
        pairs = np.ascontiguousarray(pairs, PAIR_DTYPE)
        n = len(pairs)
        dp = al.device_alloc(pairs.nbytes)
        al.h2d(dp, pairs)
        own = True
    try:
        choice = al.autotune(dp, n, "align" if xdrop else "sw_scores",
                             sample=sample, force=force)
    finally:
        if own:
            al.device_free(dp)
    dt = time.time() - t
    if log:
        log("autotune picked %s in %.1f s" % (choice, dt))
    return choice, dt
