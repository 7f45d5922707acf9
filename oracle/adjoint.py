"""FP64 reference of the reverse-mode derivative of the chi(omega) path.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
adjoint (th_chi_adjoint).  The product never imports it.

The reference (TITAN) has no derivatives, so there is no reference routine to
restate; this module differentiates the oracle's own forward model, the
eintshechi contraction (eintshechi.F90:83-105 imaginary axis, 143-165 real
axis) followed by the Dyson step of calculate_chi.F90:84-89, by hand:

  forward   H(k) = h0 + sum_n exp(i k.R_n) t_n (x) 1_spin
            G    = (z - H)^-1
            chi0 = (1/2pi) [sum_imag w (Gq(i,j)[a,b] G(j,i)[c,d] + conj(G(i,j) Gq(j,i)))
                             + sum_real (-i w)(Gq - Gq^H)(i,j)[a,b] G^H(j,i)[c,d]]
            chi  = (I + chi0 U)^-1 chi0
  reverse   (cotangents xhat = dL/dRe x + i dL/dIm x of a real loss L)
            chi0hat = M^-H chihat N^H,  Uhat = -chi0^H M^-H chihat chi^H
            Ghat from the contraction, Hhat = G^H Ghat G^H,
            that_n = sum_p exp(-i k.R_n) sum_s Hhat[(jb,s),(ib,s)],  h0hat = sum_p Hhat.

Parity pin: tests/test_adjoint_oracle.py checks every output of this module
against central finite differences of the C oracle (or_chi0, or_dyson, which
are pinned to the reference's fixtures) in random directions.  Dense loops,
small systems only.
"""
import numpy as np

_A = (1, 0, 1, 0)   # spin of the first G index per sigma = (+, up, dn, -): a_sigma
_C = (0, 0, 1, 1)   # c_sigma


def chi_index_maps(nA, nOrb):
    """(blk, r, c) of the contraction accumulator for each chi entry (X, Y),
    term 1 (b1) and term 2 (b2), as arrays [D, D]."""
    bs = 2 * nOrb
    D = 4 * nOrb * nOrb * nA
    X = np.arange(D)
    s = X & 3
    r = X >> 2
    nu = r % nOrb
    mu = (r // nOrb) % nOrb
    at = r // (nOrb * nOrb)
    a = np.asarray(_A)[s]
    c = np.asarray(_C)[s]
    # X rows (sigma1, i, mu, nu), Y cols (sigma2, j, gamma, xi)
    rr1 = (a[:, None] * nOrb + nu[:, None]) * bs + (a[None, :] * nOrb + mu[None, :])
    cc1 = (c[None, :] * nOrb + nu[None, :]) * bs + (c[:, None] * nOrb + mu[:, None])
    blk1 = at[:, None] * nA + at[None, :]
    rr2 = (a[None, :] * nOrb + mu[None, :]) * bs + (a[:, None] * nOrb + nu[:, None])
    cc2 = (c[:, None] * nOrb + mu[:, None]) * bs + (c[None, :] * nOrb + nu[None, :])
    blk2 = at[None, :] * nA + at[:, None]
    return (blk1, rr1, cc1), (blk2, rr2, cc2)


class ArrayModel:
    """H(k) from explicit tables: R[nhop,3], jb, ib, t[nhop,nO,nO] (rows =
    orbitals of jb), h0[d,d]."""

    def __init__(self, nA, nOrb, R, jb, ib, t, h0):
        self.nA, self.nOrb = nA, nOrb
        self.d = 2 * nOrb * nA
        self.D = 4 * nOrb * nOrb * nA
        self.R = np.asarray(R, dtype=float).reshape(-1, 3)
        self.jb = np.asarray(jb, dtype=int)
        self.ib = np.asarray(ib, dtype=int)
        self.t = np.asarray(t, dtype=complex)
        self.h0 = np.asarray(h0, dtype=complex)

    def hk(self, k):
        H = self.h0.copy()
        nO = self.nOrb
        ph = np.exp(1j * (self.R @ np.asarray(k, dtype=float)))
        for n in range(len(ph)):
            j, i = self.jb[n], self.ib[n]
            for s in range(2):
                r0, c0 = (2 * j + s) * nO, (2 * i + s) * nO
                H[r0:r0 + nO, c0:c0 + nO] += self.t[n] * ph[n]
        return H

    def green(self, z, k):
        return np.linalg.inv(z * np.eye(self.d) - self.hk(k))


def _pairs(model, q, omega, Ef, eta, etap, imag, real_k, real_e):
    """Yields (axis, w, Gq, G, k) for every (k,E) pair of the forward."""
    q = np.asarray(q, dtype=float)
    kim, yim, wim = imag
    for p in range(len(yim)):
        k = np.asarray(kim[p], dtype=float)
        Gq = model.green(Ef + omega + 1j * (yim[p] + eta), k + q)
        G = model.green(Ef + 1j * (yim[p] + etap), k)
        yield 0, wim[p], Gq, G, k
    if real_k is not None and abs(omega) >= 1e-15:
        kre, wre = real_k
        x2, p2 = real_e
        for e in range(len(x2)):
            for ik in range(len(wre)):
                k = np.asarray(kre[ik], dtype=float)
                Gq = model.green(x2[e] + omega + 1j * eta, k + q)
                G = model.green(x2[e] + 1j * etap, k)
                yield 1, p2[e] * wre[ik], Gq, G, k


def _blocks(M, nA, bs):
    """vec of atom blocks: out[i*nA+j] = M(i,j) flattened row-major."""
    return M.reshape(nA, bs, nA, bs).transpose(0, 2, 1, 3).reshape(nA * nA, bs * bs)


def _unblocks(V, nA, bs):
    return V.reshape(nA, nA, bs, bs).transpose(0, 2, 1, 3).reshape(nA * bs, nA * bs)


def _swap(V, nA):
    """block (i,j) <- block (j,i)"""
    return V.reshape(nA, nA, -1).transpose(1, 0, 2).reshape(nA * nA, -1)


def _btrans(V, nA, bs):
    """each flattened block transposed"""
    return V.reshape(nA * nA, bs, bs).transpose(0, 2, 1).reshape(nA * nA, bs * bs)


def _panels(ax, w, Gq, G, nA, bs):
    """A_p, B_p of one pair: imaginary axis A(i,j) = w Gq(i,j), B(i,j) = G(j,i);
    real axis A = -i w (Gq - Gq^H)(i,j), B(i,j) = G^H(j,i)."""
    if ax == 0:
        return _blocks(w * Gq, nA, bs), _swap(_blocks(G, nA, bs), nA)
    A = _blocks(-1j * w * (Gq - Gq.conj().T), nA, bs)
    return A, _btrans(np.conj(_blocks(G, nA, bs)), nA, bs)


def chi0(model, q, omega, Ef, eta, etap, imag, real_k=None, real_e=None):
    """Forward chi0 through the contraction form (numpy)."""
    nA, bs = model.nA, 2 * model.nOrb
    M = bs * bs
    T = np.zeros((2, nA * nA, M, M), dtype=complex)
    for ax, w, Gq, G, k in _pairs(model, q, omega, Ef, eta, etap, imag, real_k, real_e):
        A, B = _panels(ax, w, Gq, G, nA, bs)
        T[ax] += A[:, :, None] * B[:, None, :]
    (b1, r1, c1), (b2, r2, c2) = chi_index_maps(nA, model.nOrb)
    return (T[0][b1, r1, c1] + np.conj(T[0][b2, r2, c2]) + T[1][b1, r1, c1]) / (2 * np.pi)


def dyson(chi0_, U):
    D = chi0_.shape[0]
    return np.linalg.solve(np.eye(D) + chi0_ @ U, chi0_)


def adjoint(model, q, omega, Ef, eta, etap, imag, real_k, real_e, chi0_, chi_hat, U=None, chi=None):
    """(t_hat[nhop,nO,nO], h0_hat[d,d], U_hat[D,D] or None) for the cotangent
    chi_hat of chi (U given) or of chi0 (U None)."""
    nA, nO = model.nA, model.nOrb
    bs, d, D = 2 * nO, model.d, model.D
    M = bs * bs
    U_hat = None
    if U is not None:
        Mm = np.eye(D) + chi0_ @ U
        X = np.linalg.solve(Mm.conj().T, chi_hat)
        c0hat = X @ (np.eye(D) - U @ chi).conj().T
        U_hat = -chi0_.conj().T @ X @ chi.conj().T
    else:
        c0hat = chi_hat
    (b1, r1, c1), (b2, r2, c2) = chi_index_maps(nA, nO)
    That = np.zeros((2, nA * nA, M, M), dtype=complex)
    That[0][b1, r1, c1] += c0hat / (2 * np.pi)
    That[0][b2, r2, c2] += np.conj(c0hat) / (2 * np.pi)
    That[1][b1, r1, c1] += c0hat / (2 * np.pi)
    t_hat = np.zeros(model.t.shape, dtype=complex)
    h0_hat = np.zeros((d, d), dtype=complex)
    q = np.asarray(q, dtype=float)

    for ax, w, Gq, G, k in _pairs(model, q, omega, Ef, eta, etap, imag, real_k, real_e):
        Th = That[ax]
        A, B = _panels(ax, w, Gq, G, nA, bs)
        Ahat = np.einsum('brc,bc->br', Th, np.conj(B))
        Bhat = np.einsum('brc,br->bc', Th, np.conj(A))
        Ah = _unblocks(Ahat, nA, bs)          # cotangent of the A matrix (d x d)
        if ax == 0:
            Gq_hat = w * Ah
            G_hat = _unblocks(_swap(Bhat, nA), nA, bs)
        else:
            Gq_hat = 1j * w * (Ah + Ah.conj().T)
            G_hat = np.conj(_unblocks(_btrans(Bhat, nA, bs), nA, bs))
        for Gm, Gh, kk in ((Gq, Gq_hat, k + q), (G, G_hat, k)):
            Hh = Gm.conj().T @ Gh @ Gm.conj().T
            h0_hat += Hh
            ph = np.exp(1j * (model.R @ kk))
            for n in range(len(ph)):
                j, i = model.jb[n], model.ib[n]
                blk = (Hh[2 * j * nO:(2 * j + 1) * nO, 2 * i * nO:(2 * i + 1) * nO] +
                       Hh[(2 * j + 1) * nO:(2 * j + 2) * nO, (2 * i + 1) * nO:(2 * i + 2) * nO])
                t_hat[n] += np.conj(ph[n]) * blk
    return t_hat, h0_hat, U_hat
