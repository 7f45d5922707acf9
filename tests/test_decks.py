"""TITAN's decks and self-consistency files (titan_amd/decks.py).

tests/golden/decks/ holds data files of two of the reference's CI examples,
copied unchanged: examples/Fe_bulk_susc/{input, basis, Fe} and
examples/nanoribbon/{input, basis, A2, A3}, each with its
results_correct/TSOC/selfconsistency/selfconsistency_*.dat.  The reader must
give the same TBModel parameters as the JSON fixtures the GPU tests use (those
reproduce the reference's chi files at 1e-5), and the self-consistency writer
must reproduce the reference's files byte for byte.
"""
import glob
import json
import os

import numpy as np
import pytest

from titan_amd import decks as D

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")
CASES = (("Fe_bulk_susc", "fe_bulk_susc.json"), ("nanoribbon", "nanoribbon.json"))


@pytest.mark.parametrize("deck,fixture", CASES)
def test_deck_matches_fixture(deck, fixture):
    d = D.load_deck(os.path.join(DECKS, deck))
    with open(os.path.join(HERE, "golden", fixture)) as fh:
        ref = json.load(fh)
    for k, v in ref["system"].items():
        if k == "name" or k.startswith("element_"):
            continue
        if k == "types":   # the reader adds each element file's own lattice
            for got, want in zip(d["system"][k], v):
                assert {kk: got[kk] for kk in want} == want
            continue
        assert d["system"][k] == v, k
    for k, v in ref["run"].items():
        assert d["run"][k] == v, k


def test_fe_element_file():
    el = D.read_element(os.path.join(DECKS, "Fe_bulk_susc", "Fe"), 3)
    with open(os.path.join(HERE, "golden", "fe_bulk_susc.json")) as fh:
        ref = json.load(fh)["system"]
    assert el["fermi_level"] == ref["element_fermi_level"]
    assert el["a0"] == ref["element_a0"]
    assert el["bravais"] == ref["element_bravais"]
    assert el["U"] == [0.0, 0.07353] and el["lambda"] == [0.0, 0.004]
    assert el["occupations"] == [0.58, 0.37, 7.05] and el["sysdim"] == 3
    with pytest.raises(ValueError):
        D.read_element(os.path.join(DECKS, "Fe_bulk_susc", "Fe"), 4)


def test_ribbon_shells_and_cartesian_positions():
    b = D.read_basis(os.path.join(DECKS, "nanoribbon", "basis"))
    assert b["elements"] == ["A2", "A3"] and b["counts"] == [5, 5] and b["coord"] == "C"
    assert b["pos"][1] == pytest.approx([0.5 * 0.535, 0.8660254037844388 * 0.535, 0.0])
    el = D.read_element(os.path.join(DECKS, "nanoribbon", "A2"), 1)
    # p line with three values (fourth word is not the line id), one value on the others
    assert el["onsite"][1:4] == [-0.15, -0.15, -1000.0]
    assert el["onsite"][0] == -1000.0 and el["onsite"][4:7] == [1000.0] * 3
    assert el["U"] == [0.0, 0.0]                        # one value: Un = Um
    assert D.orbitals("px py") == [2, 3] and D.orbitals(None) == list(range(1, 10))
    fe = D.read_basis(os.path.join(DECKS, "Fe_bulk_susc", "basis"))
    assert fe["coord"] == "L" and fe["pos"] == [[0.0, 0.0, 0.0]]


@pytest.mark.parametrize("deck,norb", (("Fe_bulk_susc", 9), ("nanoribbon", 2)))
def test_sc_file_round_trip_bytes(deck, norb, tmp_path):
    src = glob.glob(os.path.join(DECKS, deck, "results_correct", "TSOC", "selfconsistency", "*.dat"))[0]
    with open(src) as fh:
        nA = sum(1 for l in fh if l.strip() and not l.startswith("!")) - 1
    sc = D.read_sc_results(src, [norb] * nA)
    dst = str(tmp_path / "sc.dat")
    D.write_sc_results(dst, sc["rho"], sc["mx"], sc["my"], sc["mz"], sc["Ef"], sc["delta_sc"])
    with open(src, "rb") as a, open(dst, "rb") as b:
        assert a.read() == b.read()


def test_sc_records_may_wrap(tmp_path):
    rng = np.random.default_rng(1)
    v = rng.standard_normal((2, 5, 3))
    p = tmp_path / "wrapped.dat"
    with open(p, "w") as fh:
        for i in range(2):
            flat = ["%.15e" % x for x in v[i].ravel()]
            fh.write(" ".join(flat[:7]) + "\n" + " ".join(flat[7:]) + " 99.0\n")   # extra word ignored
        fh.write("0.25 ! Ef\n")
    sc = D.read_sc_results(str(p), [3, 3])
    assert np.allclose(sc["mz"], v[:, 3]) and np.allclose(sc["delta_sc"], v[:, 4]) and sc["Ef"] == 0.25
    with pytest.raises(ValueError):
        D.read_sc_results(str(p), [3, 3, 3])


def test_deck_without_sc_file(tmp_path):
    """No self-consistency file: zero moments, E_F from the input key 'Ef'
    (nanoribbon) or from the element file (Fe); SOC = F switches the
    spin-orbit scale off as readElementFile does."""
    import shutil
    fe = tmp_path / "fe"
    fe.mkdir()
    for f in ("input", "basis", "Fe"):
        shutil.copy(os.path.join(DECKS, "Fe_bulk_susc", f), fe / f)
    d = D.load_deck(str(fe))
    assert d["sc_file"] is None and d["system"]["Ef"] == 0.7415
    assert d["system"]["mz"] == [[0.0] * 9]
    text = (fe / "input").read_text().replace("-> SOC = T", "-> SOC = F")
    (fe / "input").write_text(text)
    assert D.load_deck(str(fe))["system"]["socscale"] == 0.0
    rib = tmp_path / "rib"
    rib.mkdir()
    for f in ("input", "basis", "A2", "A3"):
        shutil.copy(os.path.join(DECKS, "nanoribbon", f), rib / f)
    r = D.load_deck(str(rib))
    assert r["system"]["Ef"] == -0.25 and len(r["system"]["rho"]) == 10


def test_adaptive_mesh_sizes():
    """Per-energy meshes of the imaginary axis (adaptiveMesh.F90:199-218):
    N(y) = max(ceil(N / (y/y1)^sqrt(d)), minimumBZmesh) rounded up to a cube,
    y1 the smallest y; weights of each energy sum to its contour weight; the
    deck's meshes (nkpt = minimumBZmesh = 1000) are all 10^3."""
    import math
    from titan_amd import contour as Ct
    from titan_amd.system import TBModel
    d = D.load_deck(os.path.join(DECKS, "Fe_bulk_susc"))
    m = TBModel(d["system"])
    run = dict(d["run"])
    (K, Y, W), (kp, wk) = D.work_lists(m, run)
    y, wy = Ct.imag_epoints(run["parts"], run["n1gl"])
    assert len(Y) == 1000 * len(y) and len(wk) == 1000
    run.update({"nkpt": 4096, "nk": (16, 16, 16), "minimumBZmesh": 64})
    (K, Y, W), _ = D.work_lists(m, run)
    for e in range(len(y)):
        sel = Y == y[e]
        n = max(math.ceil(4096 / (y[e] / y[0]) ** math.sqrt(3)), 64)
        edge = math.ceil(n ** (1 / 3) - 1e-9)
        assert sel.sum() == edge ** 3
        assert W[sel].sum() == pytest.approx(wy[e], rel=1e-12)
    assert (Y == y[0]).sum() == 4096 and (Y == y[-1]).sum() == 64
