"""The Newton Jacobian of the self-consistency equations, read from the static
chi0 (titan_amd.selfconsistency.jacobian), against central finite
differences of the equations themselves, both evaluated with the FP64
oracle: expectation values by the oracle's restatement of
expectation_values_greenfunction, chi0 by its restatement of eintshechi.
Un != 0 and Um != 0 (occupation, moment and Fermi-level unknowns), one and
two atoms."""
import numpy as np
import pytest

from oracle import expectation as OE
from oracle import oracle as O
from titan_amd import configs as CF
from titan_amd import selfconsistency as SC
from titan_amd.system import TBModel


def _setup(nA):
    spec = CF.bcc_supercell(nA, seed=11) if nA > 1 else CF.bcc_1site(seed=11)
    for ty in spec["types"]:
        ty["U"] = [0.3 / 13.6, 0.9 / 13.6]        # Un, Um
    run = {"nk": (2, 2, 2), "parts": 1, "n1gl": 6, "eta": 2e-2}
    return spec, run


def _F(spec, model, names, x, K, Y, W, run, n_total):
    s = SC.set_x(spec, model, names, x)
    osys = O.OracleSystem(s)
    ev = OE.expectation_values(osys, K, Y, W, s["Ef"], run["eta"])
    return SC.residual(s, TBModel(s), names, x, ev, n_total)


@pytest.mark.parametrize("nA", [1, 2])
def test_jacobian_from_chi0_matches_finite_differences(nA):
    spec, run = _setup(nA)
    model = TBModel(spec)
    names = SC.unknowns(spec, model)
    assert len(names) == nA * 6 + 1
    K, Y, W, kp, wk = SC.imag_list(model, run)
    x = SC.get_x(spec, model, names)
    n_total = 8.0 * nA
    s = SC.set_x(spec, model, names, x)
    chi0 = O.OracleSystem(s).chi0([0, 0, 0], 0.0, run["eta"], run["eta"], (K, Y, W))
    J = SC.jacobian(s, TBModel(s), names, chi0)
    h = 1e-5
    Jfd = np.zeros_like(J)
    for k in range(len(x)):
        e = np.zeros(len(x))
        e[k] = h
        Jfd[:, k] = (_F(spec, model, names, x + e, K, Y, W, run, n_total) -
                     _F(spec, model, names, x - e, K, Y, W, run, n_total)) / (2 * h)
    err = np.abs(J - Jfd).max() / np.abs(Jfd).max()
    print("nA", nA, "Jacobian vs central differences: %.2e" % err)
    assert err < 1e-7


def test_planted_sign_error_is_caught():
    """The same check fails for a Jacobian built with the spin-up and
    spin-down rows of the response swapped (a plausible index mistake)."""
    spec, run = _setup(1)
    model = TBModel(spec)
    names = SC.unknowns(spec, model)
    K, Y, W, kp, wk = SC.imag_list(model, run)
    x = SC.get_x(spec, model, names)
    s = SC.set_x(spec, model, names, x)
    chi0 = O.OracleSystem(s).chi0([0, 0, 0], 0.0, run["eta"], run["eta"], (K, Y, W))
    nO = model.nOrb
    bad = chi0.copy()
    for mu in range(nO):
        a, b = 4 * ((0 * nO + mu) * nO + mu) + 1, 4 * ((0 * nO + mu) * nO + mu) + 2
        bad[[a, b], :] = bad[[b, a], :]
    J = SC.jacobian(s, TBModel(s), names, chi0)
    Jb = SC.jacobian(s, TBModel(s), names, bad)
    assert np.abs(J - Jb).max() > 1e-3 * np.abs(J).max()
