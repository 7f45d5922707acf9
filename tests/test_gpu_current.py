"""GPU spin-current response (th_current_response, eintshe) against the
oracle's restatement of the reference's 16 spin-block terms: imaginary axis
only and with the real axis, q = 0 and q != 0, with and without an RPA
prefactor, the elimination and the spectral Green functions, one, three,
four and eight atoms (d = 18 to 144; the larger through the three-stage
form, which also runs on the small systems against the one-kernel form)."""
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


@pytest.mark.parametrize("which,q,omega", [
    ("bcc4", (0.0, 0.0, 0.0), 0.004),
    ("slab8", (0.03, 0.0, 0.0), 0.0),
    ("slab8", (0.0, 0.0, 0.0), 0.004),
])
def test_current_response_large_blocks_match_oracle(which, q, omega):
    """d = 72 and d = 144 (where a pair's G, Gq and products do not fit the
    LDS): the three-stage form against the oracle, with a prefactor."""
    spec = CF.bcc_supercell(4) if which == "bcc4" else CF.fcc001_slab(8)
    m, s = TBModel(spec), O.OracleSystem(spec)
    y, wy = Ct.imag_epoints(1, 4)
    kp, wk = Ct.uniform_kmesh(m.b, 2, 1, 1) if which == "bcc4" else Ct.uniform_kmesh(m.b, 2, 2, 1)
    K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
    x2, p2 = Ct.real_epoints(m.Ef, omega if omega else 0.004, 1, 3)
    eta, etap = 5e-3, 4e-3
    ue = np.array([0.6, 0.8, 0.0])
    rng = np.random.default_rng(3)
    P = np.eye(m.D) + 0.05 * (rng.normal(size=(m.D, m.D)) + 1j * rng.normal(size=(m.D, m.D)))
    h = m.handle()
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    print(which, "d", m.d, "D", m.D, "q", q, "omega", omega, flush=True)
    got = h.current_response(q, omega, m.Ef, eta, etap, ue, x2, p2, prefactor=P)
    h.close()
    want = OC.current_response(s, q, omega, eta, etap, ue, (K, Y, W), (kp, wk), (x2, p2), prefactor=P)
    print("  rel err %.2e (|ref| %.3e)" % (rel(got, want), np.linalg.norm(want)), flush=True)
    assert np.linalg.norm(want) > 0
    assert rel(got, want) < 1e-10


@pytest.mark.parametrize("which,omega", [("cfg0", 0.004), ("cfg2", 0.0), ("cfg2", 0.004)])
def test_three_stage_current_equals_one_kernel_form(which, omega):
    """The three-stage form (tuning key current_general) on systems the
    one-kernel form also takes: the same sums, so equal to 1e-13."""
    spec = CF.system(0 if which == "cfg0" else 2)
    m = TBModel(spec)
    y, wy = Ct.imag_epoints(1, 6)
    kp, wk = Ct.uniform_kmesh(m.b, 3, 3, 2)
    K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
    x2, p2 = Ct.real_epoints(m.Ef, omega if omega else 0.004, 1, 4)
    ue = np.array([0.3, -0.5, 0.81])
    out = {}
    for gen in (0, 1):
        h = m.handle()
        h.tune(current_general=gen, batch_mb=16)
        h.set_imag_points(K, Y, W)
        h.set_real_kmesh(kp, wk)
        out[gen] = h.current_response([0.01, 0.0, 0.0], omega, m.Ef, 5e-3, 4e-3, ue, x2, p2)
        h.close()
    print(which, "omega", omega, "rel diff %.2e" % rel(out[1], out[0]), flush=True)
    assert rel(out[1], out[0]) < 1e-13
