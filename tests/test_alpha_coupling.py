"""Host consumers of chi: the Gilbert damping rows (write_alpha,
mod_alpha.F90:128-221) and the couplings J, K (get_J_K_from_chi,
mod_coupling.F90:130-241).

The reference holds no output of either, so both are pinned by closed forms:
* a Landau-Lifshitz-Gilbert transverse susceptibility chi^{+-} = 1/(w0 - e -
  i alpha e) (and chi^{-+} = 1/(-w0 - e - i alpha e)) gives, from the inverse,
  alpha e / w0 in the damping column and gammaM = e / w0 (the slope files'
  purpose: alpha at resonance);
* atom blocks accumulate over atoms 1..i as in the reference (zeroed once,
  mod_alpha.F90:148-151): two identical uncoupled atoms give atom 2 twice
  atom 1's blocks, so its gamma column halves;
* chiSt = 4 chi_spin + 2 chi_SL: with chi_SL = (M - 4 chi_spin)/2 the block is
  M, whose inverse is chosen;
* (J, K) -> chi^-1 -> chi -> (J, K) round trip over FM and AFM pairs with
  unequal moments.
Values read back from the rows carry es16.9's ten significant digits.
"""
import os

import numpy as np
import pytest

from titan_amd import alpha as A
from titan_amd import coupling as Cp
from titan_amd.titan_io import read_data


def llg_schi(nA, w0, alpha, e, long=(0.7 + 0.1j, 0.4 - 0.2j)):
    """Block-diagonal 4nA x 4nA schi: per atom chi^{+-}, chi^{up up}, chi^{dn dn},
    chi^{-+} on the diagonal (index s nA + i)."""
    s = np.zeros((4 * nA, 4 * nA), dtype=complex)
    for i in range(nA):
        d = [1.0 / (w0 - e - 1j * alpha * e), long[0], long[1], 1.0 / (-w0 - e - 1j * alpha * e)]
        for k in range(4):
            s[k * nA + i, k * nA + i] = d[k]
    return s


def test_llg_damping_columns():
    w0, alpha, e, M, Um = 0.02, 0.013, 0.005, 2.1, 0.06
    schi = llg_schi(1, w0, alpha, e)
    rows, vals = A.alpha_rows(e, schi, schi, [M], [Um])
    inv = [float(x) for x in rows["chiinv"][0].split()]
    assert len(inv) == 36
    assert inv[1] == pytest.approx(alpha * e / w0, rel=2e-9)       # alpha column
    assert inv[2] == pytest.approx(e / (M * w0), rel=2e-9)          # gamma column
    assert inv[3] == pytest.approx(M * alpha, rel=2e-9)             # alpha/gamma column
    assert vals["gammaM"][0] == pytest.approx(e / w0, rel=1e-12)
    assert vals["v1"][0] == pytest.approx(alpha * e / w0, rel=1e-12)
    assert vals["v4"][0] == pytest.approx((e / w0) * schi[0, 0].imag / e * Um ** 2, rel=1e-12)
    # the cartesian inverse block: xx = -e - i alpha e, xy = i w0 (re/im, q fastest)
    assert complex(inv[4 + 2 * 5], inv[5 + 2 * 5]) == pytest.approx(-e - 1j * alpha * e, rel=2e-9)
    assert complex(inv[4 + 2 * 9], inv[5 + 2 * 9]) == pytest.approx(1j * w0, rel=2e-9)


def test_blocks_accumulate_over_atoms():
    w0, alpha, e = 0.02, 0.013, 0.005
    schi = llg_schi(2, w0, alpha, e)
    rows, vals = A.alpha_rows(e, schi, schi, [1.5, 1.5], [0.05, 0.05])
    r1 = np.array([float(x) for x in rows["chiinv"][0].split()])
    r2 = np.array([float(x) for x in rows["chiinv"][1].split()])
    assert r2[1] == pytest.approx(r1[1], rel=2e-9)           # ratio unchanged
    assert r2[2] == pytest.approx(r1[2] / 2, rel=2e-9)       # twice the block
    assert np.allclose(r2[4:], 2 * r1[4:], rtol=2e-9, atol=0)
    assert vals["gammaM"][1] == pytest.approx(vals["gammaM"][0] / 2, rel=1e-12)


def test_soc_block_inverse_entries():
    w0, alpha, e = 0.02, 0.013, 0.005
    schi = llg_schi(1, w0, alpha, e)
    rng = np.random.default_rng(3)
    N = np.eye(3) * 2 + 0.3 * (rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3)))
    Mb = np.linalg.inv(N)
    spin = A.STOC @ schi @ A.CTOS
    sl = (Mb - 4.0 * spin[1:, 1:]) / 2.0
    ls = np.eye(3) * (1 + 0.5j) + 0.1j * rng.standard_normal((3, 3))
    ll = np.eye(3) * (2 - 0.1j) + 0.1 * rng.standard_normal((3, 3))
    rows, _ = A.alpha_rows(e, schi, schi, [1.0], [0.05], ls=ls, sl=sl, ll=ll)
    st = [float(x) for x in rows["chiSt"][0].split()]
    assert st[1] == pytest.approx(e / N[1, 2].imag, rel=2e-9)
    assert st[2] == pytest.approx(-N[1, 1].imag / N[1, 2].imag, rel=2e-9)
    lsi = np.linalg.inv(ls)
    assert [float(x) for x in rows["chiLS"][0].split()][1] == pytest.approx(e / lsi[1, 2].imag, rel=2e-9)


def test_alpha_files(tmp_path):
    schi = llg_schi(2, 0.02, 0.01, 0.004)
    d = str(tmp_path / "A" / "Slope")
    paths = A.create_alpha_files(d, 2, soc=False, suffix="_test")
    assert len(paths) == 10
    for e in (0.004, 0.006):
        A.write_alpha(d, e, llg_schi(2, 0.02, 0.01, e), schi, [1.0, 2.0], [0.05, 0.06], suffix="_test")
    rows = read_data(os.path.join(d, "chiinv_2_test.dat"))
    assert len(rows) == 2 and len(rows[0]) == 36
    assert rows[1][0] == pytest.approx(0.006)
    with open(os.path.join(d, "sumrule_1_test.dat")) as fh:
        assert fh.readline().startswith("#     energy    ,  gammaM")
    with pytest.raises(FileNotFoundError):
        A.write_alpha(str(tmp_path / "none"), 0.004, schi, schi, [1.0, 2.0], [0.05, 0.06])
    with pytest.raises(ValueError):
        A.alpha_rows(0.0, schi, schi, [1.0, 2.0], [0.05, 0.06])


def test_J_K_round_trip():
    nA = 3
    mvec = np.array([[0, 0, 2.2], [0.1, 0, 0.7], [0, 0, -1.3]])
    mabs = np.linalg.norm(mvec, axis=1)
    c = Cp.RY_MEV
    J = np.array([[0, 3.1, -1.7], [2.4, 0, 0.9], [-0.6, 1.3, 0]])
    Kx, Ky = np.array([0.35, -0.2, 0.11]), np.array([0.15, -0.05, 0.11])
    rng = np.random.default_rng(7)
    X = 5 * np.eye(4 * nA) + 0.2 * (rng.standard_normal((4 * nA,) * 2) + 1j * rng.standard_normal((4 * nA,) * 2))
    jt = np.zeros(nA)
    for i in range(nA):
        for j in range(nA):
            if i == j:
                continue
            if mvec[i] @ mvec[j] > 0:      # FM: J = Re X[+i,+j] M_i M_j / 2
                X[i, j] = 2 * J[i, j] / (mabs[i] * mabs[j] * c) + 0.3j
                jt[i] += J[i, j]
            else:                          # AFM: J = -Re X[+i,-j] M_i M_j / 2
                X[i, 3 * nA + j] = -2 * J[i, j] / (mabs[i] * mabs[j] * c) - 0.2j
                jt[i] -= J[i, j]
    for i in range(nA):
        X[i, i] = 2 * (Kx[i] + Ky[i] - jt[i]) / (mabs[i] ** 2 * c)
        X[i, 3 * nA + i] = 2 * (Kx[i] - Ky[i]) / (mabs[i] ** 2 * c)
    r = Cp.get_J_K_from_chi(np.linalg.inv(X), mvec)
    assert np.allclose(r["J"], J, rtol=1e-9, atol=1e-9)
    assert np.allclose(r["Kx"], Kx, rtol=1e-9, atol=1e-9)
    assert np.allclose(r["Ky"], Ky, rtol=1e-9, atol=1e-9)
    text = "\n".join(r["log"])
    assert "J(1,3) = " in text and "(AFM alignment)" in text and "(FM alignment)" in text
    assert "-> Uniaxial anisotropy: Kx(3) = Ky(3)" in text
    assert "-> z is easy axis: Kx(2), Ky(2) < 0" in text and "-> z is hard axis: Kx(1), Ky(1) > 0" in text


def test_J_K_skips_nonmagnetic_atoms():
    mvec = np.array([[0, 0, 1.0], [0, 0, 0.0]])
    r = Cp.get_J_K_from_chi(np.eye(8) * 2.0, mvec)
    assert r["J"][0, 1] == 0 and r["J"][1, 0] == 0
    assert np.isnan(r["Kx"][1]) and np.isfinite(r["Kx"][0])
    with pytest.raises(ValueError):
        Cp.get_J_K_from_chi(np.eye(4), mvec)
