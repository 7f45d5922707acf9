"""GPU spin-current response (th_current_response, eintshe) against the
oracle's restatement of the reference's 16 spin-block terms: imaginary axis
only and with the real axis, q = 0 and q != 0, with and without an RPA
prefactor, the elimination and the spectral Green functions, one and three
atoms."""
import numpy as np
import pytest

from oracle import current as OC
from oracle import oracle as O
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("which,q,omega,spectral,pre", [
    ("cfg0", (0.0, 0.0, 0.0), 0.0, 0, False),
    ("cfg0", (0.05, 0.0, -0.02), 0.004, 0, True),
    ("cfg0", (0.0, 0.0, 0.0), 0.004, 1, False),
    ("cfg2", (0.0, 0.0, 0.0), 0.004, 1, True),
    ("cfg2", (0.02, 0.01, 0.0), 0.0, 0, False),
])
def test_current_response_matches_oracle(which, q, omega, spectral, pre):
    spec = CF.system(0 if which == "cfg0" else 2)
    m, s = TBModel(spec), O.OracleSystem(spec)
    y, wy = Ct.imag_epoints(1, 6)
    kp, wk = Ct.uniform_kmesh(m.b, 3, 2, 2)
    K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
    x2, p2 = Ct.real_epoints(m.Ef, omega if omega else 0.004, 1, 4)
    eta, etap = 5e-3, 4e-3
    ue = np.array([0.3, -0.5, 0.81])
    rng = np.random.default_rng(7)
    P = None
    if pre:
        P = np.eye(m.D) + 0.1 * (rng.normal(size=(m.D, m.D)) + 1j * rng.normal(size=(m.D, m.D)))
    h = m.handle()
    h.tune(spectral=spectral)
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    if spectral:   # eigenpairs in place (th_chi0 computes them), so the spectral kernels give G
        h.chi0(q, omega, m.Ef, eta, etap, x2, p2)
    got = h.current_response(q, omega, m.Ef, eta, etap, ue, x2, p2, prefactor=P)
    want = OC.current_response(s, q, omega, eta, etap, ue, (K, Y, W), (kp, wk), (x2, p2), prefactor=P)
    print(which, "q", q, "omega", omega, "spectral", spectral, "prefactor", pre,
          "rel err %.2e (|ref| %.3e)" % (rel(got, want), np.linalg.norm(want)), flush=True)
    assert np.linalg.norm(want) > 0
    assert rel(got, want) < 1e-10
    h.close()
