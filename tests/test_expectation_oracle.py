"""Pin of the oracle's ground-state expectation values (oracle/expectation.py)
by the spectral theorem: (1/pi) int_0^inf Re G(E_F + i(y + eta)) dy + 1/2 =
V diag(1/2 + arctan((E_F - lambda)/eta)/pi) V^H, so a fine Gauss-Legendre y
quadrature must give the occupations and moments of numpy's eigen-
decomposition of the oracle's H(k)."""
import numpy as np
import pytest

from oracle import expectation as OE
from oracle import oracle as O
from titan_amd import configs as CF
from titan_amd import contour as Ct


def _spectral(osys, kp, wk, Ef, eta):
    nA, nO = osys.nA, osys.nOrb
    Dm = np.zeros((osys.d, osys.d), dtype=complex)
    for k, w in zip(kp, wk):
        lam, V = np.linalg.eigh(osys.hk(k))
        f = 0.5 + np.arctan((Ef - lam) / eta) / np.pi
        Dm += w * (V * f) @ V.conj().T
    Dv = Dm.reshape(nA, 2, nO, nA, 2, nO)
    uu = np.array([[Dv[i, 0, m, i, 0, m].real for m in range(nO)] for i in range(nA)])
    dd = np.array([[Dv[i, 1, m, i, 1, m].real for m in range(nO)] for i in range(nA)])
    mp = np.array([[2 * Dv[i, 1, m, i, 0, m] for m in range(nO)] for i in range(nA)])
    return uu + dd, mp.real, mp.imag, uu - dd, mp


@pytest.mark.parametrize("cfg", [0, 2])
def test_expectation_values_match_spectral_theorem(cfg):
    osys = O.OracleSystem(CF.system(cfg))
    kp, wk = osys.kmesh_uniform(2, 2, 1)
    y, wy = Ct.imag_epoints(4, 64)
    K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
    eta = 0.02
    got = OE.expectation_values(osys, K, Y, W, osys.Ef, eta)
    want = _spectral(osys, kp, wk, osys.Ef, eta)
    for name, g, w in zip(("rho", "mx", "my", "mz", "mp"), got, want):
        assert np.abs(g - w).max() < 1e-7, name
    assert np.abs(want[0]).sum() > 1.0 and np.abs(want[3]).sum() > 1e-3   # occupations and a moment


def test_planted_spin_swap_is_caught():
    osys = O.OracleSystem(CF.system(0))
    kp, wk = osys.kmesh_uniform(2, 1, 1)
    y, wy = Ct.imag_epoints(4, 64)
    K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
    rho, mx, my, mz, mp = OE.expectation_values(osys, K, Y, W, osys.Ef, 0.02)
    want = _spectral(osys, kp, wk, osys.Ef, 0.02)
    # a swapped spin order (mz = n_down - n_up) fails the pin
    assert np.abs(-mz - want[3]).max() > 1e-3
