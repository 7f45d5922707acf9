"""GPU spin density (th_spin_density, calcSmunu) against the oracle, and the
sum rule (mod_sumrule.F90:10-109) on the device's own chi0 and <S>: the
static spin-rotation identity pins the whole chi0 path end to end."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import sumrule as OS
from titan_amd import binding as B
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd import sumrule as SR
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("which,spectral", [("cfg0", 1), ("cfg0", 0), ("bcc2", 1), ("cfg2", 1), ("cfg2", 0)])
def test_spin_density_and_sum_rule(which, spectral):
    spec = {"cfg0": lambda: CF.system(0), "bcc2": lambda: CF.bcc_supercell(2), "cfg2": lambda: CF.system(2)}[which]()
    m, s = TBModel(spec), O.OracleSystem(spec)
    y, wy = Ct.imag_epoints(2, 8)
    kp, wk = Ct.uniform_kmesh(m.b, 2, 2, 2)
    K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
    eta = 1e-3
    h = m.handle()
    h.tune(spectral=spectral)
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    S = h.spin_density(m.Ef, eta)
    S_ref = OS.spin_density(s, K, Y, W, m.Ef, eta)
    print(which, "spectral", spectral, "S vs oracle %.2e" % rel(S, S_ref), flush=True)
    assert rel(S, S_ref) < 1e-11
    for prec, tol in ((B.PREC_FP64, 1e-10), (B.PREC_MIXED, 1e-5)):
        chi0 = h.chi0([0.0, 0.0, 0.0], 0.0, m.Ef, eta, eta, precision=prec)
        diff, lhs, _ = SR.check(m, chi0, S)
        scale = np.abs(lhs).sum()
        print("  precision", prec, "sum rule |lhs - rhs| %.2e of %.2e" % (diff, scale), flush=True)
        assert diff < tol * scale
    h.close()
