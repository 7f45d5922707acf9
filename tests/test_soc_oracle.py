"""Pins of the oracle's spin-orbit traces and local-frame rotation
(oracle/soc.py) by what the mathematics fixes, since the reference holds no
fixture for them: the rotation pair is a proper cartesian rotation that takes
each atom's magnetisation to -z, the orbital angular momentum matrices obey the
angular momentum algebra, and the cross responses of an outer-product chi
factorise into traces of products computed here independently."""
import numpy as np
import pytest

from oracle import soc
from titan_amd.system import l_matrices


@pytest.mark.parametrize("theta,phi", [(0, 0), (30, 40), (90, 0), (123, -70), (10, 200), (180, 33)])
def test_rotation_pair_is_a_proper_rotation_to_minus_z(theta, phi):
    A = soc.rotation_matrices_chi(180.0 - theta, 180.0 + phi, True)
    B = soc.rotation_matrices_chi(180.0 - theta, 180.0 + phi, False)
    assert np.allclose(A @ B, np.eye(4), atol=1e-14)
    assert np.allclose(soc.STOC @ soc.CTOS, np.eye(4), atol=1e-15)
    M = soc.STOC @ A @ soc.CTOS                      # the rotation in cartesian (0, x, y, z)
    assert abs(M[0, 0] - 1) < 1e-14 and np.abs(M[0, 1:]).max() < 1e-14 and np.abs(M[1:, 0]).max() < 1e-14
    R = M[1:, 1:]
    assert np.abs(R.imag).max() < 1e-14
    R = R.real
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-14) and abs(np.linalg.det(R) - 1) < 1e-14
    t, p = np.radians(theta), np.radians(phi)
    m = np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)])
    assert np.allclose(R @ m, [0, 0, -1], atol=1e-14)


def test_angular_momentum_algebra():
    Lx, Ly, Lz = l_matrices()
    for a, b, c in ((Lx, Ly, Lz), (Ly, Lz, Lx), (Lz, Lx, Ly)):
        assert np.allclose(a @ b - b @ a, 1j * c, atol=1e-14)
    L2 = Lx @ Lx + Ly @ Ly + Lz @ Lz
    assert np.allclose(L2, np.diag([0, 2, 2, 2, 6, 6, 6, 6, 6]), atol=1e-14)
    for M in (Lx, Ly, Lz):
        assert np.allclose(M, M.conj().T)


def _lvec(nOrb):
    L = l_matrices()                                  # [s][mu][gama], orbitals 0..8
    return np.transpose(L, (1, 2, 0))[:nOrb, :nOrb, :]


def _cidx(nOrb, s, i, mu, nu):
    return 4 * ((i * nOrb + mu) * nOrb + nu) + s


@pytest.mark.parametrize("nA,nOrb", [(1, 9), (2, 9), (3, 4)])
def test_cross_responses_factorise_for_an_outer_product(nA, nOrb):
    """chi = u v^T: every trace is (left trace of u) x (right trace of v)."""
    rng = np.random.default_rng(nA * 10 + nOrb)
    D = 4 * nA * nOrb * nOrb
    u = rng.normal(size=D) + 1j * rng.normal(size=D)
    v = rng.normal(size=D) + 1j * rng.normal(size=D)
    chi = np.outer(u, v)
    L = _lvec(nOrb)
    U = lambda i, s: np.array([[u[_cidx(nOrb, s, i, m, g)] for g in range(nOrb)] for m in range(nOrb)])
    V = lambda j, s: np.array([[v[_cidx(nOrb, s, j, n, x)] for x in range(nOrb)] for n in range(nOrb)])
    # left factors: tr(L_s^T U) for the L side, trace of U for the S side; right likewise
    lL = lambda i, s: np.sum(L[:, :, s] * (U(i, 1) + U(i, 2)))
    lS = lambda i, c: sum(soc.STOC[c + 1, p] * np.trace(U(i, p)) for p in range(4))
    rS = lambda j, c: sum(np.trace(V(j, p)) * soc.CTOS[p, c + 1] for p in range(4))
    rL = lambda j, t: np.sum((V(j, 1) + V(j, 2)) * L[:, :, t])
    LS, SL, LL = soc.soc_traces(nA, nOrb, chi, L)
    for i in range(nA):
        for j in range(nA):
            for s in range(3):
                for t in range(3):
                    r, c = s * nA + i, t * nA + j
                    assert abs(LS[r, c] - lL(i, s) * rS(j, t)) < 1e-10 * (1 + abs(LS[r, c]))
                    assert abs(SL[r, c] - lS(i, s) * rL(j, t)) < 1e-10 * (1 + abs(SL[r, c]))
                    assert abs(LL[r, c] - lL(i, s) * rL(j, t)) < 1e-10 * (1 + abs(LL[r, c]))
    sch = soc.spin_traces(nA, nOrb, chi)
    for i in range(nA):
        for j in range(nA):
            for s in range(4):
                for t in range(4):
                    assert abs(sch[s * nA + i, t * nA + j] - np.trace(U(i, s)) * np.trace(V(j, t))) < 1e-10 * (
                        1 + abs(sch[s * nA + i, t * nA + j]))


def test_planted_transpose_is_caught():
    """The factorisation check fails for an oracle that reads L transposed."""
    nA, nOrb = 1, 9
    rng = np.random.default_rng(5)
    D = 4 * nA * nOrb * nOrb
    u = rng.normal(size=D) + 1j * rng.normal(size=D)
    v = rng.normal(size=D) + 1j * rng.normal(size=D)
    chi = np.outer(u, v)
    L = _lvec(nOrb)
    good, _, _ = soc.soc_traces(nA, nOrb, chi, L)
    bad, _, _ = soc.soc_traces(nA, nOrb, chi, np.transpose(L, (1, 0, 2)))
    assert np.abs(good - bad).max() > 1e-3 * np.abs(good).max()


def test_rotation_of_spin_traces_blockwise():
    nA = 2
    rng = np.random.default_rng(3)
    S = rng.normal(size=(4 * nA, 4 * nA)) + 1j * rng.normal(size=(4 * nA, 4 * nA))
    ang = [(30.0, 40.0), (120.0, -15.0)]
    R = soc.rotate_spin_traces(nA, S, ang)
    # rotating back with the inverse pair restores the traces
    A = [soc.rotation_matrices_chi(180.0 - t, 180.0 + p, True) for t, p in ang]
    B = [soc.rotation_matrices_chi(180.0 - t, 180.0 + p, False) for t, p in ang]
    for i in range(nA):
        for j in range(nA):
            blk = R.reshape(4, nA, 4, nA)[:, i, :, j]
            back = B[i] @ blk @ A[j]
            assert np.allclose(back, S.reshape(4, nA, 4, nA)[:, i, :, j], atol=1e-12)
