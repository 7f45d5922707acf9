"""Autotuner for the kernel shapes of the chi(omega) path.

Times th_chi0 on a sample of the caller's (k,E) work list for each candidate
of the library's tuning knobs (th_tune):

  green_variant  inversion kernel shape (0 default tiles, 2 one wave per
                 matrix, 3 two waves per matrix, where built for d)
  contract_3m    three instead of four real MFMA products per complex one
  splits         K-splits of the contraction (workgroups per output tile)
  batch_mb       panel memory per batch (pairs per launch)

and leaves the fastest setting on the handle.  The device time of each call
comes from th_last_timing (HIP events on the library's stream).  Tuning is
per handle: the best shapes depend on d = 2 nOrb nA, the number of atom
pairs and the batch sizes.  Results are bitwise identical for every
green_variant (same operations per matrix entry); splits and batch_mb change
the order of the FP64 sums of the contraction only, contract_3m its rounding.
"""
import time

import numpy as np

from titan_amd import binding as B

# inversion kernel shapes built per d (th_green.hip)
VARIANTS = {36: (0, 2), 54: (0, 2, 3), 72: (0, 3)}


def _knobs(h, base):
    """Candidate values of each knob, tuned one knob at a time."""
    s0 = base["splits"]
    return [("green_variant", list(VARIANTS.get(h.d, (0,)))),
            ("contract_3m", [0, 1]),
            ("splits", sorted({max(1, s0 // 2), s0, 2 * s0})),
            ("batch_mb", [256, 512, 1024])]


def tune(h, q, omega, Ef, eta, etap, x2=None, p2=None, precision=B.PREC_MIXED, repeat=2, log=None):
    """Coordinate search over the knobs on the handle's current work list
    (set a representative sample before calling): each knob in turn takes
    the value with the lowest device time, the others held.  Applies the
    result and returns {"best": {...}, "best_ms", "table": [(setting, ms)],
    "seconds"}."""
    t0 = time.perf_counter()
    base = h.tuning()
    setting = {"green_variant": base["green_variant"], "contract_3m": 0, "splits": base["splits"],
               "batch_mb": base["batch_mb"]}
    table = []

    def timed(st):
        h.tune(**st)
        best = np.inf
        for _ in range(repeat + 1):     # the first call warms caches and buffers
            h.chi0(q, omega, Ef, eta, etap, x2, p2, precision=precision)
            best = min(best, h.last_timing()["total_ms"])
        table.append((dict(st), best))
        if log:
            log("autotune %s %.3f ms" % (st, best))
        return best

    best_ms = timed(setting)
    for knob, values in _knobs(h, base):
        for v in values:
            if v == setting[knob]:
                continue
            trial = dict(setting, **{knob: v})
            ms = timed(trial)
            if ms < best_ms:
                best_ms, setting = ms, trial
    h.tune(**setting)
    return {"best": setting, "best_ms": best_ms, "default": base, "table": table,
            "seconds": time.perf_counter() - t0}
