"""The sum rule (mod_sumrule.F90:10-109) as a pin of the oracle's chi0 and
spin density: at omega = q = 0 with eta = eta' the static spin-rotation
identity holds point by point in (k, z), so the oracle's discrete sums satisfy
it to rounding; the host assembly of both sides (titan_amd.sumrule) is checked
with it.  A planted error (S_y with the wrong sign) breaks it."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import sumrule as OS
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd import sumrule as SR
from titan_amd.system import TBModel


def _setup(which):
    spec = {"cfg0": lambda: CF.system(0), "bcc2": lambda: CF.bcc_supercell(2)}[which]()
    m, s = TBModel(spec), O.OracleSystem(spec)
    y, wy = Ct.imag_epoints(2, 8)
    kp, wk = Ct.uniform_kmesh(m.b, 2, 2, 2)
    K = np.tile(kp, (len(y), 1))
    Y = np.repeat(y, len(wk))
    W = np.outer(wy, wk).ravel()
    return m, s, (K, Y, W), (kp, wk)


@pytest.mark.parametrize("which", ["cfg0", "bcc2"])
def test_sum_rule_holds_for_the_oracle(which):
    m, s, (K, Y, W), (kp, wk) = _setup(which)
    eta = 1e-3
    chi0 = s.chi0([0.0, 0.0, 0.0], 0.0, eta, eta, (K, Y, W), (kp, wk), (np.zeros(0), np.zeros(0)))
    S = OS.spin_density(s, K, Y, W, m.Ef, eta)
    diff, lhs, rhs = SR.check(m, chi0, S)
    scale = np.abs(lhs).sum()
    print(which, "sum |lhs - rhs| %.3e of sum |lhs| %.3e" % (diff, scale))
    assert scale > 1e-3                    # a magnetic system: the rule is not 0 = 0
    assert diff < 1e-9 * max(1.0, scale)
    # planted mistake in the oracle: S_y with the wrong sign (ud and du swapped)
    S_bad = S.copy()
    S_bad[:, 1] = -S[:, 1]
    assert SR.check(m, chi0, S_bad)[0] > 1e3 * max(diff, 1e-12)
