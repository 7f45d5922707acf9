"""Host consumers of the electric-field response (titan_amd.disturbances) on
inputs whose answer is fixed by hand: which trace picks up which spin
component, the halving of the spin parts, the orbital moment of p_x/p_y,
the effective field through an identity susceptibility, the magnetisation
angle rule.  Their agreement with the reference's own output files is the
GPU test test_gpu_disturbances."""
import numpy as np

from titan_amd import disturbances as Dz


def _tchi(nA, nO, entries):
    t = np.zeros((4 * nA * nO * nO, 4), dtype=complex)
    for (sigma, i, mu, nu), v in entries.items():
        t[4 * ((i * nO + mu) * nO + nu) + sigma, 1] = v      # half in each used column
        t[4 * ((i * nO + mu) * nO + nu) + sigma, 2] = v
    return t


def test_charge_and_spin_components():
    nA, nO = 2, 2
    lv = [Dz.type_lvec([2, 3])] * nA
    # up on atom 0 orbital 1: charge 2v, S_z v (halved from 2v), nothing else
    d, tot, sd, _, raw = Dz.disturbances(_tchi(nA, nO, {(1, 0, 1, 1): 0.5}), nA, nO, lv)
    assert np.allclose(d[:, 0], [1.0, 0, 0, 0.5, 0, 0, 0]) and np.allclose(d[:, 1], 0)
    assert np.allclose(raw[3, 0], 1.0) and np.allclose(tot, d.sum(axis=1))
    # sdmat (halved): + = x + iy, up = 0 + z, down = 0 - z, - = x - iy
    assert np.allclose(sd, 0.5 * np.array([0, 0, 2, 0, 0, 0, 0, 0]))
    # + on atom 1: S_x and S_y = S_x / i
    d, _, _, _, _ = Dz.disturbances(_tchi(nA, nO, {(0, 1, 0, 0): 0.5}), nA, nO, lv)
    assert np.allclose(d[:, 1], [0, 0.5, -0.5j, 0, 0, 0, 0])


def test_orbital_moment_of_px_py():
    lv = Dz.type_lvec([2, 3])
    assert np.allclose(lv[0], 0) and np.allclose(lv[1], 0)
    assert np.allclose(lv[2], [[0, -1j], [1j, 0]])
    # L_z = sum lz(mu, nu) (t_up + t_down)(mu, nu): an up entry at (px, py)
    d, _, _, ld, _ = Dz.disturbances(_tchi(1, 2, {(1, 0, 0, 1): 0.5}), 1, 2, [lv])
    assert np.allclose(ld[0], [[0, 1], [0, 0]]) and np.allclose(d[6, 0], -1j)


def test_effective_field_through_identity():
    nA = 3
    rng = np.random.default_rng(0)
    sd = rng.standard_normal(4 * nA) + 1j * rng.standard_normal(4 * nA)
    b, tot = Dz.effective_field(np.eye(4 * nA), sd, nA)
    p, u, dn, m = (sd[s * nA:(s + 1) * nA] for s in range(4))
    assert np.allclose(b, [u + dn, 0.5 * (p + m), -0.5j * (p - m), 0.5 * (u - dn)])
    assert np.allclose(tot, b.sum(axis=1))


def test_magnetisation_angles_and_rotation():
    mz = [[0.3, 0.2], [1e-10, 0.0], [0.0, 0.0]]
    mx = [[0.0, 0.0], [0.0, 0.0], [0.25, 0.25]]
    my = [[0.0, 0.0], [0.0, 0.0], [0.25, 0.25]]
    m, th, ph = Dz.magnetisation_angles(mx, my, mz)
    assert np.allclose(th, [0.0, 0.0, 90.0]) and np.allclose(ph, [0.0, 0.0, 45.0])
    assert np.allclose(m[:, 2], [0.5, 0.5, 0.0])
    lv = Dz.type_lvec([2, 3, 4])
    # the frame of a moment along +x: L_z' = L_x (theta = 90, phi = 0)
    lp = Dz.rotated_lvec(lv, 90.0, 0.0)
    assert np.allclose(lp[2], lv[0]) and np.allclose(lp[0], -lv[2]) and np.allclose(lp[1], lv[1])


def test_response_rows_follow_the_reference_format(tmp_path):
    """Rows as mod_disturbances.F90:205-220 writes them: phase, real/amplitude
    and imaginary/amplitude at or below 1e-15 written as 0, nine es16.9
    fields; the row of the nanoribbon's Beff_z at site 1 reproduced."""
    from titan_amd import titan_io as IO
    z = complex(-1.326220122e-04, 1.808964766e-05)
    row = IO.response_line(1e-6, z, 0.0, 0.0)
    want = ("1.000000000E-06   1.338500410E-04  -1.326220122E-04   1.808964766E-05   3.006029201E+00"
            "   0.000000000E+00   1.351486151E-01   0.000000000E+00   0.000000000E+00")
    # the value comes from the file's rounded re/im, so the derived fields
    # agree to the 10-digit rounding; the layout exactly
    assert np.allclose([float(v) for v in row.split()], [float(v) for v in want.split()], rtol=1e-8, atol=0)
    assert [f for f in row.split(" ") if f][0] == "1.000000000E-06" and len(row) == 9 * 16 + 8 * 2
    vals = {n: (np.array([1.0 + 2.0j, 0.0]), 3.0) for _, n, _ in IO.RESPONSE_FILES}
    paths = IO.write_response_files(str(tmp_path), [0.5], [vals], [([0.0, 0.0], [0.0, 0.0])], [(0.0, 0.0)], "_x")
    assert len(paths) == 17 * 3
    r = IO.read_data(str(tmp_path / "SD" / "Sz_pos=1_x.dat"))[0]
    assert np.allclose(r[:4], [0.5, abs(1 + 2j), 1.0, 2.0])
    assert IO.read_data(str(tmp_path / "SOT" / "XCTz_total_x.dat"))[0][2] == 3.0


def test_field_direction_bases():
    from titan_amd import decks as D
    a = [[1.0, 0.0, 0.0], [0.5, np.sqrt(3) / 2, 0.0], [0.0, 0.0, 2.0]]
    assert np.allclose(D.field_direction("cartesian", ["2", "0", "0"], a), [1, 0, 0])
    assert np.allclose(D.field_direction("bravais", ["1", "1"], a), np.array([1.5, np.sqrt(3) / 2, 0]) / np.sqrt(3))
    assert np.allclose(D.field_direction("spherical", ["90", "90"], a), [0, 1, 0])
