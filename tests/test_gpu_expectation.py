"""GPU ground-state expectation values (th_expectation_values,
expectation_values_greenfunction) against the oracle, elimination and
spectral Green functions, one and three atoms."""
import numpy as np
import pytest

from oracle import expectation as OE
from oracle import oracle as O
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("cfg,spectral", [(0, 0), (0, 1), (2, 1)])
def test_expectation_values_match_oracle(cfg, spectral):
    spec = CF.system(cfg)
    m, s = TBModel(spec), O.OracleSystem(spec)
    y, wy = Ct.imag_epoints(2, 8)
    kp, wk = Ct.uniform_kmesh(m.b, 2, 2, 2)
    K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
    eta = 1e-3
    h = m.handle()
    h.tune(spectral=spectral)
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    if spectral:
        h.chi0([0.0, 0.0, 0.0], 0.0, m.Ef, eta, eta)
    got = h.expectation_values(m.Ef, eta)
    want = OE.expectation_values(s, K, Y, W, m.Ef, eta)
    for name, g, w in zip(("rho", "mx", "my", "mz", "mp"), got, want):
        err = np.abs(g - w).max()
        print("cfg", cfg, "spectral", spectral, name, "max abs err %.2e" % err, flush=True)
        assert err < 1e-11 * max(1.0, np.abs(w).max())
    h.close()
