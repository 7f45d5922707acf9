"""Autotuner for the kernel shapes of the chi(omega) path.

Times th_chi0 on a sample of the caller's (k,E) work list for each candidate
of the library's tuning knobs (th_tune):

  spectral       G(k,z) from one eigendecomposition per k (1) or by
                 elimination of every z - H (0), where built for d
  green_variant  elimination kernel shape (0 default tiles, 2 one wave per
                 matrix, 3 two waves per matrix, where built for d)
  overlap        Green functions of one batch concurrent with the contraction
                 of the previous one on a second stream (1: eliminations
                 only, 2: also the spectral kernels)
  contract_f16   mixed precision: contraction on fp16 MFMA with two-term
                 splits of the FP32 panels (1) or on FP32 MFMA (0)
  contract_tiles split-panel contraction with eight waves holding 2-D blocks
                 of the output tile (1) or seven 16-row strips (0)
  contract_3m    three instead of four real MFMA products per complex one
                 (FP32/FP64 MFMA contraction)
  splits         K-splits of the contraction (workgroups per output tile)
  batch_mb       panel memory per batch (pairs per launch)

and leaves the fastest setting on the handle.  The device time of each call
comes from th_last_timing (HIP events on the library's stream).  Tuning is
per handle: the best shapes depend on d = 2 nOrb nA, the number of atom
pairs and the batch sizes.  Results are bitwise identical for every
green_variant, overlap and contract_tiles setting (same operations per matrix entry); splits
and batch_mb change the order of the FP64 sums of the contraction only,
contract_3m and spectral their rounding.
"""
import time

import numpy as np

from titan_amd import binding as B

# elimination kernel shapes built per d, and the d with a spectral kernel
# (th_green.hip)
VARIANTS = {36: (0, 2), 54: (0, 2, 3), 72: (0, 3)}
SPECTRAL_D = (18, 36, 54, 72, 144, 288)


def _knobs(h, base, setting):
    """Candidate values of the next knob given the setting so far (knobs are
    tuned one at a time; the elimination knobs only matter without the
    spectral path)."""
    s0 = base["splits"]
    yield "spectral", [1, 0] if h.d in SPECTRAL_D else [0]
    if not setting["spectral"]:
        yield "green_variant", list(VARIANTS.get(h.d, (0,)))
        yield "overlap", [1, 0]
    else:
        yield "overlap", [1, 2]
    yield "contract_f16", [1, 0]
    if setting["contract_f16"]:
        yield "contract_tiles", [1, 0]
    yield "contract_3m", [0, 1]
    yield "splits", sorted({max(1, s0 // 2), s0, 2 * s0})
    yield "batch_mb", [256, 512, 1024]


def tune(h, q, omega, Ef, eta, etap, x2=None, p2=None, precision=B.PREC_MIXED, repeat=2, log=None):
    """Coordinate search over the knobs on the handle's current work list
    (set a representative sample before calling): each knob in turn takes
    the value with the lowest device time, the others held.  Applies the
    result and returns {"best": {...}, "best_ms", "table": [(setting, ms)],
    "seconds"}."""
    t0 = time.perf_counter()
    base = h.tuning()
    setting = {"spectral": 1 if h.d in SPECTRAL_D else 0, "green_variant": base["green_variant"],
               "overlap": 1, "contract_f16": 1, "contract_tiles": 1, "contract_3m": 0, "splits": base["splits"],
               "batch_mb": base["batch_mb"]}
    table = []

    def timed(st):
        h.tune(**st)
        best = np.inf
        try:
            for _ in range(repeat + 1):     # the first call warms caches and buffers
                h.chi0(q, omega, Ef, eta, etap, x2, p2, precision=precision)
                best = min(best, h.last_timing()["total_ms"])
        except B.TitanHipError as err:      # a shape this system cannot take (e.g. a batch too large)
            if log:
                log("autotune %s rejected: %s" % (st, err))
            best = np.inf
        table.append((dict(st), best))
        if log:
            log("autotune %s %.3f ms" % (st, best))
        return best

    best_ms = timed(setting)
    for knob, values in _knobs(h, base, setting):
        for v in values:
            if v == setting[knob]:
                continue
            trial = dict(setting, **{knob: v})
            ms = timed(trial)
            if ms < best_ms:          # in place: _knobs reads the live setting
                best_ms = ms
                setting.update(trial)
    h.tune(**setting)
    if not np.isfinite(best_ms):
        raise B.TitanHipError("autotune: no setting ran on this system")
    return {"best": setting, "best_ms": best_ms, "default": base, "table": table,
            "seconds": time.perf_counter() - t0}
