"""TITAN's susceptibility file rows and comparator (titan_amd/titan_io.py):
the writer reproduces the reference's own Fe result files byte for byte from
their numbers, the comparator accepts them and catches a planted change
(also of a negative value, which the reference's one-sided relative test
lets through), and the eigen-decomposition of chi^{+-} is LAPACK's."""
import os

import numpy as np

from titan_amd import titan_io as IO

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _schi_rows(path):
    """(energies, q, [4x4 schi]) from a one-atom chi file."""
    rows = IO.read_data(path)
    es, qs, S = [], [], []
    for r in rows:
        es.append(r[0])
        qs.append(r[1])
        z = np.array(r[2::2]) + 1j * np.array(r[3::2])
        S.append(z.reshape(4, 4).T)        # n fastest within m: element [n, m]
    return es, qs, S


def test_writer_reproduces_reference_files_exactly():
    for kind in ("chi", "chihf"):
        path = os.path.join(GOLD, "fe_%s_1_1.dat" % kind)
        text = open(path).read().splitlines()
        es, qs, S = _schi_rows(path)
        assert text[0] == IO.HEADER[kind]
        for line, e, q, s in zip(text[1:], es, qs, S):
            assert IO.chi_line(e, q, s, 0, 0, 1) == line


def test_comparator(tmp_path):
    path = os.path.join(GOLD, "fe_chi_1_1.dat")
    assert IO.compare_files(path, path) == []
    es, qs, S = _schi_rows(path)
    for target in (0, 1):   # the negative real part of chi+- and a positive entry
        S2 = [s.copy() for s in S]
        if target == 0:
            S2[0][0, 0] = S2[0][0, 0] * (1 + 1e-5)      # negative value
            assert S2[0][0, 0].real < 0
        else:
            S2[2][0, 0] = complex(S2[2][0, 0].real, S2[2][0, 0].imag * (1 + 1e-5))
        p = IO.write_chi_files(str(tmp_path / str(target)), "chi", es, qs[0], S2, 1)[0]
        assert len(IO.compare_files(p, path, 1e-7)) == 1


def test_diagonalize_is_an_eigendecomposition():
    rng = np.random.default_rng(2)
    nA = 5
    schi = rng.normal(size=(4 * nA, 4 * nA)) + 1j * rng.normal(size=(4 * nA, 4 * nA))
    w, V = IO.diagonalize_susceptibilities(schi, nA)
    M = schi[:nA, :nA]
    assert np.allclose(M @ V, V * w, atol=1e-12)
    assert np.allclose(np.linalg.norm(V, axis=0), 1.0)
    assert len(IO.eval_line(0.1, 0.0, w).split()) == 2 + 2 * nA
    assert len(IO.evec_line(0.1, 0.0, V, 2).split()) == 2 + 2 * nA
