"""Uneven batches: many cells of different sizes through the chi(omega) path
at once.

Each cell is one th_handle with its own HIP stream.  The cells' frequency
sweeps are issued from a pool of host threads (the ctypes calls release the
GIL), so the kernels of cells of different sizes run concurrently on the
GPU and small cells fill the gaps of large ones.  Tile shapes are tuned once
per distinct matrix size d on a sample of the first cell of that size
(titan_amd.autotune) and applied to every cell of that size.  Every cell's
buffers are bounded: panels by batch_mb per handle, the omega-independent
caches (panel cache, eigenpairs) by cache_mb per handle, so that 40 cells
fit next to each other in HBM.  By default a cell's panel batch grows with
its size (64 MB per site, at most 1 GiB): a 16-site cell's chi0 partial sums
(0.46 GB per launch) are then written once per ~800 pairs instead of once
per ~190.
"""
import concurrent.futures as cf
import time

import numpy as np

from titan_amd import autotune as AT
from titan_amd import binding as B
from titan_amd import contour as Ct
from titan_amd.system import TBModel


def cell_batch_mb(nA):
    """Default panel batch of a cell of nA sites (MB)."""
    return int(min(1024, max(64, 64 * nA)))


class Cell:
    def __init__(self, spec, run, kmesh=None, batch_mb=None, cache_mb=4096):
        self.spec, self.run = spec, run
        self.model = m = TBModel(spec)
        self.y, self.wy = Ct.imag_epoints(run["parts"], run["n1gl"])
        nk = tuple(kmesh if kmesh is not None else run["nk"])
        self.kp, self.wk = Ct.uniform_kmesh(m.b, *nk)
        self.h = m.handle()
        self.h.tune(batch_mb=cell_batch_mb(m.nA) if batch_mb is None else batch_mb, cache_mb=cache_mb)
        self.set_mesh(self.kp, self.wk)

    def set_mesh(self, kp, wk):
        self.h.set_imag_points(np.tile(kp, (len(self.y), 1)), np.repeat(self.y, len(wk)),
                               np.outer(self.wy, wk).ravel())
        self.h.set_real_kmesh(kp, wk)

    def pairs(self, omegas):
        om = np.asarray(omegas, dtype=float)
        nx = self.run["parts3"] * self.run["n3gl"]
        return int(len(self.wk) * (len(self.y) * len(om) + nx * np.sum(np.abs(om) >= 1e-15)))

    def sweep(self, omegas, precision):
        r = self.run
        return self.h.sweep(r["q"], omegas, self.model.Ef, r["eta"], r["etap"], r["parts3"], r["n3gl"],
                            self.model.Umat, precision=precision)

    def close(self):
        self.h.close()


class UnevenBatch:
    """cells = [(spec, run)]; kmesh overrides every cell's k-mesh (tests)."""

    def __init__(self, cells, kmesh=None, batch_mb=None, cache_mb=4096):
        self.cells = [Cell(s, r, kmesh, batch_mb, cache_mb) for s, r in cells]

    def autotune(self, precision=B.PREC_MIXED, sample=8, log=None):
        """One tuning per distinct d, on every sample-th k-point of the first
        cell of that size; the best setting goes to all cells of that size."""
        best = {}
        for c in self.cells:
            d = c.model.d
            if d in best:
                c.h.tune(**best[d]["best"])
                continue
            kp, wk = c.kp[::sample], c.wk[::sample] * sample
            c.set_mesh(kp, wk)
            om = float(c.run["omega"][-1])
            x2, p2 = Ct.real_epoints(c.model.Ef, om, c.run["parts3"], c.run["n3gl"])
            best[d] = AT.tune(c.h, c.run["q"], om, c.model.Ef, c.run["eta"], c.run["etap"], x2, p2,
                              precision=precision, repeat=1, log=log)
            c.set_mesh(c.kp, c.wk)
        return best

    def sweep(self, precision=B.PREC_MIXED, omegas=None, threads=8):
        """Runs every cell's sweep concurrently; returns (results, seconds,
        pairs), results[i] = (schi0, schi, cond) of cell i."""
        t0 = time.perf_counter()
        with cf.ThreadPoolExecutor(max_workers=threads) as ex:
            futs = [ex.submit(c.sweep, c.run["omega"] if omegas is None else omegas, precision)
                    for c in self.cells]
            res = [f.result() for f in futs]
        dt = time.perf_counter() - t0
        pairs = sum(c.pairs(c.run["omega"] if omegas is None else omegas) for c in self.cells)
        return res, dt, pairs

    def close(self):
        for c in self.cells:
            c.close()
