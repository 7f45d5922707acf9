"""Autotuner for the kernel shapes of the chi(omega) path.

Times th_chi0 on a sample of the caller's (k,E) work list for each candidate
of the library's tuning knobs (th_tune):

  green_variant  inversion kernel shape (0 default tiles, 2 one wave per
                 matrix, 3 two waves per matrix, where built for d)
  splits         K-splits of the contraction (workgroups per output tile)
  batch_mb       panel memory per batch (pairs per launch)

and leaves the fastest setting on the handle.  The device time of each call
comes from th_last_timing (HIP events on the library's stream).  Tuning is
per handle: the best shapes depend on d = 2 nOrb nA, the number of atom
pairs and the batch sizes.  Results are bitwise identical for every
green_variant (same operations per matrix entry); splits and batch_mb change
the order of the FP64 sums of the contraction only.
"""
import itertools
import time

import numpy as np

from titan_amd import binding as B

# inversion kernel shapes built per d (th_green.hip)
VARIANTS = {36: (0, 2), 54: (0, 2, 3), 72: (0, 3)}


def candidates(h, splits0):
    variants = VARIANTS.get(h.d, (0,))
    splits = sorted({max(1, splits0 // 2), splits0, 2 * splits0})
    return list(itertools.product(variants, splits, (256, 512, 1024)))


def tune(h, q, omega, Ef, eta, etap, x2=None, p2=None, precision=B.PREC_MIXED, repeat=2, log=None):
    """Times every candidate on the handle's current work list (set a
    representative sample before calling) and applies the fastest.
    Returns {"best": {...}, "table": [(setting, ms), ...], "seconds": s}."""
    t0 = time.perf_counter()
    base = h.tuning()
    table = []
    for v, s, b in candidates(h, base["splits"]):
        h.tune(green_variant=v, splits=s, batch_mb=b)
        best = np.inf
        for _ in range(repeat + 1):     # first call warms caches and buffers
            h.chi0(q, omega, Ef, eta, etap, x2, p2, precision=precision)
            best = min(best, h.last_timing()["total_ms"])
        setting = {"green_variant": v, "splits": s, "batch_mb": b}
        table.append((setting, best))
        if log:
            log("autotune %s %.3f ms" % (setting, best))
    setting, ms = min(table, key=lambda t: t[1])
    h.tune(**setting)
    return {"best": setting, "best_ms": ms, "default": base, "table": table,
            "seconds": time.perf_counter() - t0}
