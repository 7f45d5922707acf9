"""calculate_chi-style driver over the C ABI (calculate_chi.F90:2-242 for the
omega loop, eintshechi.F90:2-178 for chi0, calculate_chi.F90:84-107 for
the Dyson step and the orbital trace)."""
import numpy as np

from titan_amd import binding as B
from titan_amd import contour as Ct
from titan_amd.system import TBModel

# algorithmic FLOPs per (k,E) pair (real flops, complex MAC = 8):
#   imaginary-axis pair: 2 * (4 nOrb^2)^2 * 8 * nA^2, real-axis pair: half of it
#   one inversion: 8 d^3, Dyson: 24 D^3 per omega


def contraction_flops_imag(nA, nOrb=9):
    return 2 * (4 * nOrb * nOrb) ** 2 * 8 * nA * nA


def contraction_flops_real(nA, nOrb=9):
    return (4 * nOrb * nOrb) ** 2 * 8 * nA * nA


class ChiWorkload:
    """A seeded (or fixture) system on the device with its (k,E) work list."""

    def __init__(self, spec, run):
        self.spec, self.run = spec, run
        self.model = TBModel(spec)
        m = self.model
        self.y, self.wy = Ct.imag_epoints(run["parts"], run["n1gl"])
        nk = tuple(run["nk"])
        self.kp, self.wk = Ct.uniform_kmesh(m.b, *nk)
        self.nk = len(self.wk)
        self.n_x = run["parts3"] * run["n3gl"]
        self.h = m.handle()
        self.h.set_imag_points(np.tile(self.kp, (len(self.y), 1)), np.repeat(self.y, self.nk),
                               np.outer(self.wy, self.wk).ravel())
        self.h.set_real_kmesh(self.kp, self.wk)

    def pairs(self, omegas):
        """(k,E)-points of a sweep: sum over omega of N_k (N_y + [omega != 0] N_x)."""
        om = np.asarray(omegas, dtype=float)
        return int(self.nk * (len(self.y) * len(om) + self.n_x * np.sum(np.abs(om) >= 1e-15)))

    def flops(self, omegas):
        """Algorithmic FLOPs of a sweep (contraction + inversions + Dyson)."""
        m = self.model
        om = np.asarray(omegas, dtype=float)
        nz = int(np.sum(np.abs(om) >= 1e-15))
        inv = 8.0 * m.d ** 3
        ci, cr = contraction_flops_imag(m.nA, m.nOrb), contraction_flops_real(m.nA, m.nOrb)
        F = len(om) * self.nk * len(self.y) * (ci + inv) + nz * self.nk * self.n_x * (cr + 2 * inv)
        F += self.nk * len(self.y) * inv + len(om) * 24.0 * m.D ** 3
        return F

    def contraction_flops(self, omegas):
        m = self.model
        om = np.asarray(omegas, dtype=float)
        nz = int(np.sum(np.abs(om) >= 1e-15))
        return (len(om) * self.nk * len(self.y) * contraction_flops_imag(m.nA, m.nOrb)
                + nz * self.nk * self.n_x * contraction_flops_real(m.nA, m.nOrb))

    def sweep(self, omegas, precision=B.PREC_FP64, full=False):
        r = self.run
        return self.h.sweep(r["q"], omegas, self.model.Ef, r["eta"], r["etap"], r["parts3"], r["n3gl"],
                            self.model.Umat, precision=precision, full=full)

    def close(self):
        self.h.close()
