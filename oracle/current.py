"""FP64 numpy restatement of the parallel spin-current response eintshe
(eintshe.F90:2-224) and of dtdksub (dtdksub.F90:2-28) on the C oracle's
Green functions.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
th_current_response.  The product (titan_amd) never imports it.

Per (k, E) pair eintshe accumulates, for every chi index X = (sigma, i, mu, nu)
(sigmaimunu2i, 4 ((i nOrb + mu) nOrb + nu) + sigma) and current channel
sigma' = 1..4, the 16 spin-block terms written out at eintshe.F90:130-149
(imaginary axis) and :240-258 (real axis) with the velocity vertex
dtdk = sum_n i (u.R_n) e^{i k.R_n} t_n (u the field direction), then
tchiorbiikl = prefactor (sum_p w_p df_p) / 2 pi (the prefactor product is
linear, so it is applied once to the sum instead of per pair).

Pins (tests/test_current_oracle.py): dtdk is the directional derivative of
the hopping part of H(k) (central finite differences of the oracle's H);
with the vertex replaced by the identity and k + q = k, the two spin channels
sigma' = 2, 3 together reduce every term to a product of two resolvents of
one H, fixed in closed form by (z1 - H)^-1 (z2 - H)^-1 =
((z2 - H)^-1 - (z1 - H)^-1) / (z1 - z2).
"""
import numpy as np

# block of G: rows (i, s), columns (j, s'); u = 0, d = 1
_S = {"u": 0, "d": 1}

# (Gq block, G block, G block, Gq block) of the two imaginary-axis products,
# eintshe.F90:130-149: [gfAB(nu,gama,i,j,1) gfCD(xi,mu,l,i,2) +
# conjg(gfEF(mu,xi,i,l,2) gfGH(gama,nu,j,i,1))] dtdk(gama,xi,j,l)
IMAG_TABLE = {
    (1, 1): ("dd", "uu", "uu", "dd"), (1, 2): ("du", "uu", "uu", "ud"),
    (1, 3): ("dd", "du", "ud", "dd"), (1, 4): ("du", "du", "ud", "ud"),
    (2, 1): ("ud", "uu", "uu", "du"), (2, 2): ("uu", "uu", "uu", "uu"),
    (2, 3): ("ud", "du", "ud", "du"), (2, 4): ("uu", "du", "ud", "uu"),
    (3, 1): ("dd", "ud", "du", "dd"), (3, 2): ("du", "ud", "du", "ud"),
    (3, 3): ("dd", "dd", "dd", "dd"), (3, 4): ("du", "dd", "dd", "ud"),
    (4, 1): ("ud", "ud", "du", "du"), (4, 2): ("uu", "ud", "du", "uu"),
    (4, 3): ("ud", "dd", "dd", "du"), (4, 4): ("uu", "dd", "dd", "uu"),
}
# (Gq block, Gq block, G block) of the real-axis term, eintshe.F90:240-258:
# -i (gfAB(nu,gama,i,j,1) - conjg(gfCD(gama,nu,j,i,1))) conjg(gfEF(mu,xi,i,l,2)) dtdk(gama,xi,j,l)
REAL_TABLE = {
    (1, 1): ("dd", "dd", "uu"), (1, 2): ("du", "ud", "uu"), (1, 3): ("dd", "dd", "ud"), (1, 4): ("du", "ud", "ud"),
    (2, 1): ("ud", "du", "uu"), (2, 2): ("uu", "uu", "uu"), (2, 3): ("ud", "du", "ud"), (2, 4): ("uu", "uu", "ud"),
    (3, 1): ("dd", "dd", "du"), (3, 2): ("du", "ud", "du"), (3, 3): ("dd", "dd", "dd"), (3, 4): ("du", "ud", "dd"),
    (4, 1): ("ud", "du", "du"), (4, 2): ("uu", "uu", "du"), (4, 3): ("ud", "du", "dd"), (4, 4): ("uu", "uu", "dd"),
}


def blocks(G, nA, nO):
    """gf[s][s'][i, j, a, b] = G[(i, s, a), (j, s', b)] (TITAN's gfuu/gfud/gfdu/gfdd)."""
    B = np.asarray(G).reshape(nA, 2, nO, nA, 2, nO)
    return {a + b: B[:, _S[a], :, :, _S[b], :].transpose(0, 2, 1, 3) for a in "ud" for b in "ud"}


def dtdk(osys, k, ue):
    """dtdk[j, l, gama, xi] = sum over hops n from l to j of t_n[gama, xi] i (u.R_n) e^{i k.R_n}
    (dtdksub.F90:14-26; rows = orbitals of the hop's first atom, as H(k))."""
    R, jb, ib, t, _h0, _U, _b = osys.tables()
    nA, nO = osys.nA, osys.nOrb
    T = np.zeros((nA, nA, nO, nO), dtype=complex)
    ph = 1j * (R @ np.asarray(ue, dtype=float)) * np.exp(1j * (R @ np.asarray(k, dtype=float)))
    for n in range(len(R)):
        T[jb[n], ib[n]] += t[n] * ph[n]
    return T


def pair_df(Gq, G, T, nA, nO, real):
    """df[X, sigma'] of one pair (eintshe.F90:130-149 imaginary, :240-258 real),
    as the reference writes the 16 terms, with the vertex T[j, l, gama, xi]."""
    q, g = blocks(Gq, nA, nO), blocks(G, nA, nO)
    df = np.zeros((4 * nA * nO * nO, 4), dtype=complex)
    for (s, sp), tab in (REAL_TABLE if real else IMAG_TABLE).items():
        if real:
            A, C, E = tab
            # -i (q_A(nu,gama,i,j) - conj q_C(gama,nu,j,i)) conj g_E(mu,xi,i,l) T(gama,xi,j,l)
            M = q[A] - np.conj(q[C].transpose(1, 0, 3, 2))          # [i, j, nu, gama]
            v = -1j * np.einsum("ijng,jlgx,ilmx->imn", M, T, np.conj(g[E]))
        else:
            A, Cb, E, Gb = tab
            t1 = np.einsum("ijng,jlgx,lixm->imn", q[A], T, g[Cb])
            t2 = np.conj(np.einsum("ilmx,jlgx,jign->imn", g[E], np.conj(T), q[Gb]))
            v = t1 + t2
        # X = 4 ((i nO + mu) nO + nu) + sigma
        df[(s - 1)::4, sp - 1] += v.reshape(-1)
    return df


def current_response(osys, q, omega, eta, etap, ue, imag_points, real_k=None, real_e=None, prefactor=None):
    """tchiorbiikl[D, 4] of eintshe(q, omega) for the field direction ue:
    imag_points = (k[n,3], y[n], w[n]); real_k = (k[m,3], w[m]); real_e =
    (x2[nx], p2[nx]); prefactor D x D (None: identity, the HF response)."""
    nA, nO, Ef = osys.nA, osys.nOrb, osys.Ef
    q = np.asarray(q, dtype=float)
    acc = np.zeros((4 * nA * nO * nO, 4), dtype=complex)
    K, Y, W = imag_points
    for k, y, w in zip(K, Y, W):
        Gq = osys.green(Ef + omega, y + eta, np.asarray(k) + q)
        G = osys.green(Ef, y + etap, k)
        acc += w * pair_df(Gq, G, dtdk(osys, k, ue), nA, nO, False)
    if abs(omega) >= 1e-15 and real_k is not None:
        kr, wr = real_k
        x2, p2 = real_e
        for ie in range(len(x2)):
            for ik in range(len(wr)):
                Gq = osys.green(x2[ie] + omega, eta, np.asarray(kr[ik]) + q)
                G = osys.green(x2[ie], etap, kr[ik])
                acc += p2[ie] * wr[ik] * pair_df(Gq, G, dtdk(osys, kr[ik], ue), nA, nO, True)
    if prefactor is not None:
        acc = np.asarray(prefactor) @ acc
    return acc / (2.0 * np.pi)
