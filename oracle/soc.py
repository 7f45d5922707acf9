"""FP64 numpy restatement of the post-processing of one susceptibility:
orbital traces with the spin-orbit cross responses and the rotation of the
spin traces to the local frames (calculate_chi.F90:91-185).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of
th_chi_traces.  The product (titan_amd) never imports it.

Conventions (as the reference's):
  chi[X, Y]  D x D complex, X = cidx(sigma, i, mu, nu) = 4 ((i nOrb + mu) nOrb + nu)
             + sigma, sigma = 0..3 for (+, up, down, -) (mod_System.F90:158-223)
  traces     index sigmai2i(s, i) = s nA + i (mod_System.F90:171)
  lvec       lvec[mu, gama, s] = <mu|L_s|gama>, s = x, y, z (mod_magnet.F90:295-358)
  STOC, CTOS spin (+, up, down, -) <-> cartesian (0, x, y, z) (mod_constants.F90:34-45)

Pins (tests/test_soc_oracle.py): the rotation pair satisfies A B = I,
StoC A CtoS = diag(1, R) with R a proper rotation, and R m(theta, phi) = -z
for the reference's angle convention; the cross responses factorise exactly
for an outer-product chi (traces of products computed independently); L
obeys [Lx, Ly] = i Lz and L^2 = l(l+1) on the s, p, d shells.
"""
import numpy as np

# Fortran reshape(..., [4, 4]) fills column by column (mod_constants.F90:34-45)
STOC = np.array([[0, 0.5, -0.5j, 0], [1, 0, 0, 0.5], [1, 0, 0, -0.5], [0, 0.5, 0.5j, 0]], dtype=complex).T
CTOS = np.array([[0, 0.5, 0.5, 0], [1, 0, 0, 1], [1j, 0, 0, -1j], [0, 1, -1, 0]], dtype=complex).T


def rotation_matrices_chi(theta, phi, left):
    """Left (A) or right (B) 4x4 rotation of a spin-spin susceptibility in the
    (+, up, down, -) basis, angles in degrees (mod_rotation_matrices.F90:6-76)."""
    t, p = np.radians(theta), np.radians(phi)
    e, em = np.exp(1j * p), np.exp(-1j * p)
    a, b, s = 1.0 + np.cos(t), 1.0 - np.cos(t), np.sin(t)
    if left:
        R = np.array([[em * a, -s, s, -e * b],
                      [em * s, a, b, e * s],
                      [-em * s, b, a, -e * s],
                      [-em * b, -s, s, e * a]])
    else:
        R = np.array([[e * a, e * s, -e * s, -e * b],
                      [-s, a, b, -s],
                      [s, b, a, s],
                      [-em * b, em * s, -em * s, em * a]])
    return 0.5 * R


def _blocks(nA, nOrb, chi):
    """chi as X[i, mu, nu, sigma, j, nu', xi', sigma']."""
    return np.asarray(chi).reshape(nA, nOrb, nOrb, 4, nA, nOrb, nOrb, 4)


def spin_traces(nA, nOrb, chi):
    """schi[s nA + i, s' nA + j] = sum_{mu, nu} chi[(s, i, mu, mu), (s', j, nu, nu)]
    (calculate_chi.F90:94-107)."""
    X = _blocks(nA, nOrb, chi)
    d1 = np.einsum("imms jnnt -> isjt".replace(" ", ""), X)
    return d1.transpose(1, 0, 3, 2).reshape(4 * nA, 4 * nA)


def soc_traces(nA, nOrb, chi, lvec):
    """<<L, S>>, <<S, L>>, <<L, L>> in cartesian components, each 3nA x 3nA with
    index s nA + i (calculate_chi.F90:109-140).  lvec is the atoms' type's
    (all atoms share it here, as the reference reads atom i's type also for
    the orbitals of j)."""
    X = _blocks(nA, nOrb, chi)
    L = np.asarray(lvec)                                   # [mu, gama, s]
    ud = X[:, :, :, 1, ...] + X[:, :, :, 2, ...]           # spin up + down on the left
    # LS: sum_{mu,gama,nu,p} L[mu,gama,s] (X[i mu gama (1|2); j nu nu p]) CtoS[p, s'+1]
    diag_r = np.einsum("imgjnnp->imgjp", ud)
    LS = np.einsum("mgs,imgjp,pt->isjt", L, diag_r, CTOS[:, 1:])
    # LL: sum L[mu,gama,s] (X[(1|2)] (1|2)) L[nu,xi,s']
    udud = ud[..., 1] + ud[..., 2]                         # [i, mu, gama, j, nu, xi]
    LL = np.einsum("mgs,imgjnx,nxt->isjt", L, udud, L)
    # SL: sum StoC[s+1, p] (X[i mu mu p; j nu xi (1|2)]) L[nu, xi, s']
    Y = X[..., 1] + X[..., 2]                              # [i, mu, mu', p, j, nu, xi]
    diag_l = np.einsum("immpjnx->ipjnx", Y)
    SL = np.einsum("sp,ipjnx,nxt->isjt", STOC[1:, :], diag_l, L)
    shape = lambda M: M.transpose(1, 0, 3, 2).reshape(3 * nA, 3 * nA)
    return shape(LS), shape(SL), shape(LL)


def rotate_spin_traces(nA, schi, angles):
    """Spin traces in the local frames: block (i, j) -> A_i block B_j with A_i,
    B_i the left/right matrices at (180 - theta_i, 180 + phi_i)
    (calculate_chi.F90:143-185; angles[i] = (theta_i, phi_i) in degrees of the
    magnetisation of atom i)."""
    S = np.asarray(schi).reshape(4, nA, 4, nA)
    out = np.empty_like(S)
    A = [rotation_matrices_chi(180.0 - t, 180.0 + p, True) for t, p in angles]
    B = [rotation_matrices_chi(180.0 - t, 180.0 + p, False) for t, p in angles]
    for i in range(nA):
        for j in range(nA):
            out[:, i, :, j] = A[i] @ S[:, i, :, j] @ B[j]
    return out.reshape(4 * nA, 4 * nA)
