"""FP64 numpy restatement of calcSmunu (mod_sumrule.F90:162-245) on the C
oracle's Green functions: <S^{mu nu}_ii> = sum_p w_p (G + G^H)/2 pi over
z = E_F + i(y_p + eta), + 1/2 on the diagonal of the up-up and down-down blocks.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
th_spin_density.  Pinned by the sum rule itself (tests/test_sumrule_oracle.py):
with the oracle's chi0 at omega = q = 0 and eta = eta' the identity holds to
rounding, which a wrong block, sign or weight would break.
"""
import numpy as np


def spin_density(osys, K, Y, W, Ef, eta):
    """S[i, k, mu, nu] (k = x, y, z) of the imaginary list (K, Y, W)."""
    nA, nO = osys.nA, osys.nOrb
    uu = np.zeros((nA, nO, nO), dtype=complex)
    dd, ud, du = uu.copy(), uu.copy(), uu.copy()
    for k, y, w in zip(K, Y, W):
        G = osys.green(Ef, y + eta, k)
        for i in range(nA):
            up = slice(2 * i * nO, (2 * i + 1) * nO)
            dn = slice((2 * i + 1) * nO, (2 * i + 2) * nO)
            # block[mu, nu] += G[nu s, mu s'] + conj(G[mu s', nu s])
            uu[i] += (G[up, up].T + G[up, up].conj()) * w
            dd[i] += (G[dn, dn].T + G[dn, dn].conj()) * w
            ud[i] += (G[up, dn].T + G[dn, up].conj()) * w
            du[i] += (G[dn, up].T + G[up, dn].conj()) * w
    f = 1.0 / (2.0 * np.pi)
    uu, dd, ud, du = uu * f, dd * f, ud * f, du * f
    for i in range(nA):
        uu[i] += 0.5 * np.eye(nO)
        dd[i] += 0.5 * np.eye(nO)
    return np.stack([ud + du, 1j * (ud - du), uu - dd], axis=1)
