"""Pins of the oracle's spin-current response (oracle/current.py) by what the
mathematics fixes, since the reference holds no fixture for it: the velocity
vertex is the directional derivative of H(k), and with the vertex replaced by
the identity at k + q = k the spin channels 2 and 3 of every term reduce to
products of two resolvents of one Hamiltonian, which the resolvent identity
gives in closed form."""
import numpy as np
import pytest

from oracle import current as CU
from oracle import oracle as O
from titan_amd import configs as CF

# sigma -> (spin of the nu row of Gq, spin of the mu column of G)
SIG = {1: (1, 0), 2: (0, 0), 3: (1, 1), 4: (0, 1)}


@pytest.fixture(scope="module", params=[0, 2])
def osys(request):
    return O.OracleSystem(CF.system(request.param))


def test_vertex_is_the_derivative_of_h(osys):
    nO = osys.nOrb
    for k, ue in (([0.1, 0.2, 0.3], [1.0, 0.0, 0.0]), ([-0.3, 0.05, 0.4], [0.2, -0.5, 0.84])):
        k, ue = np.array(k), np.array(ue)
        T = CU.dtdk(osys, k, ue)
        h = 1e-5
        D = (osys.hk(k + h * ue) - osys.hk(k - h * ue)) / (2 * h)
        Dv = D.reshape(osys.nA, 2, nO, osys.nA, 2, nO)
        for s in range(2):
            assert np.abs(Dv[:, s, :, :, s, :].transpose(0, 2, 1, 3) - T).max() < 1e-8 * (1 + np.abs(T).max())
        assert np.abs(Dv[:, 0, :, :, 1, :]).max() < 1e-9


def _closed_form(osys, real):
    """(pair_df with the identity vertex, the expected channels 2 + 3)."""
    nA, nO, d = osys.nA, osys.nOrb, osys.d
    k = np.array([0.13, -0.21, 0.34])
    z1, z2 = osys.Ef + 0.004 + 0.02j, osys.Ef + 0.03j
    if real:
        z1, z2 = osys.Ef + 0.004 + 0.001j, osys.Ef - 0.002 + 0.0015j
    G1 = osys.green(z1.real, z1.imag, k)
    G2 = osys.green(z2.real, z2.imag, k)
    I = np.zeros((nA, nA, nO, nO), dtype=complex)
    for i in range(nA):
        I[i, i] = np.eye(nO)
    df = CU.pair_df(G1, G2, I, nA, nO, real)
    if real:
        # -i (G1 - G1^H) G2^H, G(z)^H = G(conj z) for Hermitian H
        Em = -1j * ((G2.conj().T - G1) / (z1 - np.conj(z2)) - (G2.conj().T - G1.conj().T) / (np.conj(z1) - np.conj(z2)))
    else:
        R = (G2 - G1) / (z1 - z2)
        Em = None
    want = np.zeros((4 * nA * nO * nO,), dtype=complex)
    for s, (rn, cm) in SIG.items():
        for i in range(nA):
            for mu in range(nO):
                for nu in range(nO):
                    r, c = (2 * i + rn) * nO + nu, (2 * i + cm) * nO + mu
                    X = 4 * ((i * nO + mu) * nO + nu) + s - 1
                    want[X] = Em[r, c] if real else R[r, c] + np.conj(R[c, r])
    return df[:, 1] + df[:, 2], want


@pytest.mark.parametrize("real", [False, True])
def test_identity_vertex_reduces_to_resolvent_identity(osys, real):
    got, want = _closed_form(osys, real)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_planted_spin_block_swap_is_caught(monkeypatch):
    osys = O.OracleSystem(CF.system(0))
    tab = dict(CU.IMAG_TABLE)
    tab[(2, 2)] = ("uu", "ud", "uu", "uu")       # G block of the first product moved
    monkeypatch.setattr(CU, "IMAG_TABLE", tab)
    got, want = _closed_form(osys, False)
    assert np.abs(got - want).max() > 1e-3 * np.abs(want).max()


def test_response_is_linear_in_prefactor_and_weights():
    osys = O.OracleSystem(CF.system(0))
    rng = np.random.default_rng(4)
    K = rng.uniform(-0.5, 0.5, size=(3, 3))
    Y, W = np.array([0.05, 0.2, 1.0]), np.array([0.3, 0.2, 0.5])
    ue = [0.0, 0.0, 1.0]
    a = CU.current_response(osys, [0.01, 0, 0], 0.0, 5e-4, 5e-4, ue, (K, Y, W))
    b = CU.current_response(osys, [0.01, 0, 0], 0.0, 5e-4, 5e-4, ue, (K, Y, 2 * W))
    P = rng.normal(size=(osys.D, osys.D)) + 1j * rng.normal(size=(osys.D, osys.D))
    c = CU.current_response(osys, [0.01, 0, 0], 0.0, 5e-4, 5e-4, ue, (K, Y, W), prefactor=P)
    assert np.allclose(b, 2 * a, rtol=1e-12, atol=1e-14)
    assert np.allclose(c, P @ a, rtol=1e-10, atol=1e-12)
