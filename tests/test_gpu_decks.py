"""TITAN's own examples end to end from their deck files on the GPU: the run
directory (input card, basis, element files, self-consistency file, copied
unchanged from the reference's examples into tests/golden/decks/) is read by
titan_amd.decks, the chi(omega) sweep runs on the device, and the results are
written as TITAN's susceptibility files and checked by the reference's
comparator against its own result files (Fe_bulk_susc), or entry by entry
against the reference's numbers (nanoribbon, 100 chi files per kind)."""
import os

import numpy as np
import pytest

from titan_amd import contour as Ct
from titan_amd import decks as D
from titan_amd import titan_io as IO
from titan_amd.system import TBModel

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


@pytest.mark.gpu
def test_fe_bulk_susc_deck_end_to_end(tmp_path):
    deck = D.load_deck(os.path.join(DECKS, "Fe_bulk_susc"))
    m, run = TBModel(deck["system"]), deck["run"]
    (K, Y, W), (kp, wk) = D.work_lists(m, run)     # minimumBZmesh = nkpt here: one mesh
    h = m.handle()
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    oms = Ct.omega_grid(run["emin"], run["emax"], run["nEner"])
    print("Fe deck: d", m.d, "k-points", len(wk), "energies", len(oms), flush=True)
    s0, sc, cond = h.sweep(run["q"], oms, m.Ef, run["eta"], run["etap"], run["parts3"], run["n3gl"], m.Umat)
    h.close()
    out = tmp_path / "results" / "TSOC" / "1Sites"
    for kind, sub, rows in (("chi", "RPA", sc), ("chihf", "HF", s0)):
        p = IO.write_chi_files(str(out / sub), kind, oms, 0.0, rows, 1)[0]
        diffs = IO.compare_files(p, os.path.join(HERE, "golden", "fe_%s_1_1.dat" % kind), 1e-5)
        print(kind, "differences at 1e-5:", diffs[:3], flush=True)
        assert diffs == []


@pytest.mark.gpu
def test_nanoribbon_deck_end_to_end(ribbon_fixture):
    from tests.conftest import ribbon_blocks
    deck = D.load_deck(os.path.join(DECKS, "nanoribbon"))
    m, run = TBModel(deck["system"]), deck["run"]
    (K, Y, W), (kp, wk) = D.work_lists(m, run)
    h = m.handle()
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    print("nanoribbon deck: d", m.d, "k-points", len(wk), flush=True)
    s0, sc, cond = h.sweep(run["q"], [run["emin"]], m.Ef, run["eta"], run["etap"], run["parts3"], run["n3gl"],
                           m.Umat)
    h.close()
    n = 0
    for kind, ours in (("chihf", s0[0]), ("chi", sc[0])):
        for blk, ref in ribbon_blocks(ours, ribbon_fixture, kind):
            big = np.abs(ref) > 1e-9
            if big.any():
                assert np.max(np.abs(blk[big] - ref[big]) / np.abs(ref[big])) < 1e-9
                n += 1
    assert n == 200


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [0, 1])
def test_adaptive_imaginary_mesh_matches_oracle(prec):
    """minimumBZmesh below nkpt: each imaginary energy gets its own mesh
    (generateAdaptiveMeshes, adaptiveMesh.F90:23-74, 199-218), from 6^3
    k-points at the smallest y down to 3^3; chi0 on the device against the
    oracle on the same (E, k) list, FP64 within 1e-10, mixed within 1e-6."""
    from oracle import oracle as O
    from titan_amd import configs as CF
    spec = CF.bcc_1site()
    m, s = TBModel(spec), O.OracleSystem(spec)
    run = {"nkpt": 216, "nk": (6, 6, 6), "minimumBZmesh": 27, "parts": 2, "n1gl": 8}
    (K, Y, W), (kp, wk) = D.work_lists(m, run)
    assert len(set(np.round(Y, 14))) == 16 and len(Y) < 16 * 216
    om = 2e-3
    x2, p2 = Ct.real_epoints(m.Ef, om, 1, 6)
    h = m.handle()
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    print("adaptive mesh: %d imaginary pairs, precision %d" % (len(Y), prec), flush=True)
    g = h.chi0([0.0, 0.0, 0.0], om, m.Ef, 5e-4, 5e-4, x2, p2, precision=prec)
    h.close()
    o = s.chi0([0.0, 0.0, 0.0], om, 5e-4, 5e-4, (K, Y, W), (kp, wk), (x2, p2))
    err = np.linalg.norm(g - o) / np.linalg.norm(o)
    assert err < (1e-10 if prec == 0 else 1e-6), err
