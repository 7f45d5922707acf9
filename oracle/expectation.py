"""FP64 numpy restatement of expectation_values_greenfunction
(mod_expectation.F90:28-149, without superconductivity) on the C oracle's
Green functions.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
th_expectation_values.  Pinned (tests/test_expectation_oracle.py) by the
spectral theorem: the imaginary-axis integral of G(E_F + i(y + eta)) is the
density matrix V diag(1/2 + arctan((E_F - lambda)/eta)/pi) V^H of every
k-point, so with a fine y quadrature the oracle must reproduce occupations and
moments computed from numpy's eigen-decomposition of the oracle's H(k).
"""
import numpy as np


def expectation_values(osys, K, Y, W, Ef, eta):
    """(rho, mx, my, mz, mp), each [nA, nOrb], for the imaginary list (K, Y, W)."""
    nA, nO = osys.nA, osys.nOrb
    imguu = np.zeros((nA, nO))
    imgdd = np.zeros((nA, nO))
    gud = np.zeros((nA, nO), dtype=complex)
    gdu = np.zeros((nA, nO), dtype=complex)
    for k, y, w in zip(K, Y, W):
        G = osys.green(Ef, y + eta, k)
        for i in range(nA):
            for mu in range(nO):
                u, d = (2 * i) * nO + mu, (2 * i + 1) * nO + mu
                gud[i, mu] += G[u, d] * w                 # gf(mu, mup, i, i)
                gdu[i, mu] += G[d, u] * w                 # gf(mup, mu, i, i)
                imguu[i, mu] += G[u, u].real * w
                imgdd[i, mu] += G[d, d].real * w
    imguu = imguu / np.pi + 0.5
    imgdd = imgdd / np.pi + 0.5
    mp = (gdu + np.conj(gud)) / np.pi
    return imguu + imgdd, mp.real, mp.imag, imguu - imgdd, mp
