"""The reference's consumers of eintshe's response (calculate_all.F90:272-411)
fed with the device's tchiorbiikl on the reference's nanoribbon example
(itype 8: ten sites, p_x/p_y with SOC, field along x, one frequency
omega = 1e-6 Ry, U = 0): charge, spin and orbital disturbances, effective
fields and torques per site and in total against the example's own result
files (results_correct/TSOC/10Sites/{CD,SD,LD,Beff,SOT}, copied into
tests/golden/nanoribbon.json by make_nanoribbon.py).  Every quantity is
held to 1e-8 of the largest entry of its family (the files carry ten
significant digits; the quantities that vanish by symmetry, S_x, S_y, L_x,
L_y, the torques, must vanish to that level too)."""
import numpy as np
import pytest

from titan_amd import contour as Ct
from titan_amd import disturbances as Dz
from titan_amd import titan_io as IO
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu

FAMILIES = {"Cd": ("Cd",), "S": ("Sx", "Sy", "Sz"), "L": ("Lx", "Ly", "Lz"),
            "Beff": ("Beff0", "Beffx", "Beffy", "Beffz"), "SOT": ("SOTx", "SOTy", "SOTz"),
            "XCT": ("XCTx", "XCTy", "XCTz")}


def test_ribbon_responses_match_reference_files(ribbon_fixture, tmp_path):
    d = ribbon_fixture
    spec, run = d["system"], d["run"]
    m = TBModel(spec)
    nA, nO = m.nA, m.nOrb
    y, wy = Ct.imag_epoints(run["parts"], run["n1gl"])
    kp, wk = Ct.uniform_kmesh(m.b, run["nkpt"], 1, 1)
    om = run["emin"]
    x2, p2 = Ct.real_epoints(m.Ef, om, run["parts3"], run["n3gl"])
    h = m.handle()
    h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
    h.set_real_kmesh(kp, wk)
    print("ribbon: d", m.d, "D", m.D, "k-points", len(wk), flush=True)
    chi0 = h.chi0(run["q"], om, m.Ef, run["eta"], run["etap"], x2, p2)
    # prefactor (I + chi_hf U)^-1 (calculate_all.F90:117-120)
    pre = np.linalg.inv(np.eye(m.D) + chi0 @ m.Umat)
    _, sc, _ = h.sweep(run["q"], [om], m.Ef, run["eta"], run["etap"], run["parts3"], run["n3gl"], m.Umat)
    ue = np.asarray(run["dirEfield"], dtype=float)
    ue = ue / np.linalg.norm(ue)
    tchi = h.current_response(run["q"], om, m.Ef, run["eta"], run["etap"], ue, x2, p2, prefactor=pre)
    h.close()
    lvec = [Dz.type_lvec(spec["orbs"]) for _ in range(nA)]
    dist, dtot, sd, _, raw = Dz.disturbances(tchi, nA, nO, lvec)
    beff, btot = Dz.effective_field(sc[0], sd, nA)
    lam = [spec["socscale"] * spec["types"][t]["lambda"][0] for t in spec["type"]]
    lamd = [spec["socscale"] * spec["types"][t]["lambda"][1] for t in spec["type"]]
    um = [spec["types"][t]["U"][1] for t in spec["type"]]
    mvec, th, ph = Dz.magnetisation_angles(spec["mx"], spec["my"], spec["mz"])
    lpvec = [Dz.rotated_lvec(lvec[i], th[i], ph[i]) for i in range(nA)]
    T, Ttot = Dz.torques(tchi, nA, nO, spec["orbs"], lpvec, lam, lamd, um, mvec, raw)
    ours = {"Cd": (dist[0], dtot[0]), "Sx": (dist[1], dtot[1]), "Sy": (dist[2], dtot[2]), "Sz": (dist[3], dtot[3]),
            "Lx": (dist[4], dtot[4]), "Ly": (dist[5], dtot[5]), "Lz": (dist[6], dtot[6]),
            "Beff0": (beff[0], btot[0]), "Beffx": (beff[1], btot[1]), "Beffy": (beff[2], btot[2]),
            "Beffz": (beff[3], btot[3]),
            "SOTx": (T[0, 0], Ttot[0, 0]), "SOTy": (T[0, 1], Ttot[0, 1]), "SOTz": (T[0, 2], Ttot[0, 2]),
            "XCTx": (T[1, 0], Ttot[1, 0]), "XCTy": (T[1, 1], Ttot[1, 1]), "XCTz": (T[1, 2], Ttot[1, 2])}
    ref = d["responses"]
    # the files as the reference writes them, read back by the comparator's reader
    vals = Dz.response_values(dist, dtot, beff, btot, T, Ttot)
    paths = IO.write_response_files(str(tmp_path), [om], [vals], [(th, ph)], [(0.0, 0.0)])
    assert len(paths) == 17 * (nA + 1)
    for k in ("Cd", "Sz", "Lz", "Beffz"):
        row = IO.read_data(str(tmp_path / ("SD" if k[0] == "S" else "LD" if k[0] == "L" else
                                           "CD" if k == "Cd" else "Beff") / ("%s_pos=1.dat" % k)))[0]
        assert np.allclose(row[2:4], ref[k]["pos"][0], rtol=1e-8, atol=1e-8 * np.abs(ref[k]["pos"][0]).max())
    checked = 0
    for fam, kinds in FAMILIES.items():
        refs = {k: (np.array([complex(*v) for v in ref[k]["pos"]]), complex(*ref[k]["total"])) for k in kinds}
        scale = max(max(np.abs(r[0]).max(), abs(r[1])) for r in refs.values())
        for k in kinds:
            got_pos, got_tot = ours[k]
            err = max(np.abs(got_pos - refs[k][0]).max(), abs(got_tot - refs[k][1]))
            print("%-6s family scale %.3e  max |ours - reference| %.3e" % (k, scale, err), flush=True)
            assert err <= 1e-8 * max(scale, 1e-300) or (scale == 0.0 and err < 1e-14), k
            checked += 1
    assert checked == 17
