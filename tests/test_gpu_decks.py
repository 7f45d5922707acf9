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
    y, wy = Ct.imag_epoints(run["parts"], run["n1gl"])
    n = round(run["nkpt"] ** (1 / 3))
    kp, wk = Ct.uniform_kmesh(m.b, n, n, n)
    h = m.handle()
    h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
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
    y, wy = Ct.imag_epoints(run["parts"], run["n1gl"])
    kp, wk = Ct.uniform_kmesh(m.b, run["nkpt"], 1, 1)
    h = m.handle()
    h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
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
