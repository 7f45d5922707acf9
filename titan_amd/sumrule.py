"""Sum rule check of the static susceptibility (mod_sumrule.F90:10-160).

At omega = 0 and q = 0 spin-rotation invariance of everything in H but the
local field sigma.B_eff (exchange and spin-orbit terms) gives, per atom i and
orbital pair (mu, nu),

    eps_{mnk} <S^k_{mu nu}>_i = 2 sum_{j,gamma,xi,q,p} eps_{nqp}
                                chi_cart[(m,i,mu,nu),(p,j,gamma,xi)] B^q_eff(j,gamma,xi)

with chi the Hartree-Fock susceptibility chi0 in cartesian spin components.
The check passes when sum |lhs - rhs| < 1e-5 (mod_sumrule.F90:98).  The
identity holds point by point in (k, z) when chi0 and <S> use the same
imaginary-axis points with eta = eta', so the discrete sums satisfy it to
rounding.  <S> comes from th_spin_density, chi0 from th_chi0 (both on the
device); this module assembles B_eff and both sides on the host, as the
reference does.
"""
import numpy as np

from titan_amd.system import l_matrices

# spin (+, up, down, -) <-> cartesian (0, x, y, z) (mod_constants.F90:34-45)
STOC = np.array([[0, 1, 1, 0], [0.5, 0, 0, 0.5], [-0.5j, 0, 0, 0.5j], [0, 0.5, -0.5, 0]], dtype=complex)
CTOS = np.array([[0, 1, 1j, 0], [0.5, 0, 0, 1], [0.5, 0, 0, -1], [0, 1, -1j, 0]], dtype=complex)


def levi_civita():
    e = np.zeros((3, 3, 3))
    for a, b, c in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        e[a, b, c], e[a, c, b] = 1.0, -1.0
    return e


def magnetisation_d(model):
    """(mx, my, mz) of the d orbitals of every atom, [nA, 3] (mxd, myd, mzd)."""
    spec, nA, nO = model.spec, model.nA, model.nOrb
    dm = model.orbs >= 5
    out = np.zeros((nA, 3))
    for c, key in enumerate(("mx", "my", "mz")):
        a = np.asarray(spec.get(key, np.zeros((nA, nO))), dtype=float).reshape(nA, nO)
        out[:, c] = a[:, dm].sum(axis=1)
    return out


def beff(model, mvec=None):
    """B_eff[i, mu, nu, c], c = x, y, z (Beffective, mod_sumrule.F90:112-160):
    -Um m_c / 2 on the d diagonal, plus lambda_p,d <mu|L_c|nu> / 2 on the p and d
    blocks (no external field)."""
    nA, nO = model.nA, model.nOrb
    m = magnetisation_d(model) if mvec is None else np.asarray(mvec, dtype=float)
    L = l_matrices()
    o = model.orbs - 1
    scale = float(model.spec.get("socscale", 1.0))
    B = np.zeros((nA, nO, nO, 3), dtype=complex)
    for i in range(nA):
        ty = model.spec["types"][model.type[i]]
        Um = ty["U"][1]
        lam_p, lam_d = scale * ty["lambda"][0], scale * ty["lambda"][1]
        for c in range(3):
            for mu in range(nO):
                if o[mu] >= 4:
                    B[i, mu, mu, c] += -0.5 * Um * m[i, c]
            for mu in range(nO):
                for nu in range(nO):
                    if 1 <= o[mu] <= 3 and 1 <= o[nu] <= 3:
                        B[i, mu, nu, c] += 0.5 * lam_p * L[c][o[mu], o[nu]]
                    elif o[mu] >= 4 and o[nu] >= 4:
                        B[i, mu, nu, c] += 0.5 * lam_d * L[c][o[mu], o[nu]]
    return B


def check(model, chi0, S, mvec=None):
    """Both sides of the sum rule.  chi0: D x D at omega = q = 0 (eta = eta'),
    S: [nA, 3, nOrb, nOrb] from Handle.spin_density.  Returns (sum |lhs - rhs|,
    lhs, rhs) with lhs, rhs [3, 3, nOrb, nOrb, nA] (m, n, mu, nu, i)."""
    nA, nO = model.nA, model.nOrb
    X = np.asarray(chi0).reshape(nA, nO, nO, 4, nA, nO, nO, 4)
    # cartesian components on both sides: StoC (p, r) chi (r, t) CtoS (t, q)
    Xc = np.einsum("pr,imnrjgxt,tq->imnpjgxq", STOC, X, CTOS)
    B = beff(model, mvec)
    eps = levi_civita()
    lhs = np.einsum("mnk,ikab->mnabi", eps, np.asarray(S))
    # rhs(m,n,mu,nu,i) = 2 sum eps(n,q,p) Xc[(m+1,i,mu,nu),(p+1,j,g,x)] B[j,g,x,q]
    XcB = np.einsum("iabmjgxp,jgxq->iabmpq", Xc[:, :, :, 1:, :, :, :, 1:], B)
    rhs = 2.0 * np.einsum("nqp,iabmpq->mnabi", eps, XcB)
    return float(np.sum(np.abs(lhs - rhs))), lhs, rhs
