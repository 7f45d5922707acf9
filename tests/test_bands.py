"""H(k) set-up against the reference's band-structure fixtures.

The reference's band-structure runs write the eigenvalues of h0 + H_hop(k)
along a path of k-points (band_structure, mod_band_structure.F90:74-140), so
they pin the whole host set-up that feeds the chi(omega) path:

  Nb_100_Monolayer  a (100) monolayer cut from a bcc bulk element file: the
                    integrals are scaled from the element's neighbour
                    distance (5.40 bohr) to the monolayer's (5.30 bohr) by
                    Andersen's law (TightBinding.F90:144-162), s, p and d
                    orbitals with spin-orbit coupling, and hee from the
                    self-consistency file with Un = Um != 0 and the
                    element's initial occupations (rho0);
  graphene          two atoms, pz orbitals only, honeycomb geometry.

Data files of both examples (input, basis, element file, kbands, the
self-consistency and initial-occupation files, the band-structure result)
are copied unchanged into tests/golden/decks/.  The file prints 9
significant digits (es16.8): eigenvalues must agree to 5e-8 Ry.  Without the
scaling law the Nb bands move by ~1e-2 Ry, which the same check catches."""
import glob
import os

import numpy as np
import pytest

from oracle import oracle as O
from titan_amd import decks as D
from titan_amd.system import TBModel

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")


def read_kbands(path):
    pts = {}
    with open(path) as fh:
        for line in fh:
            line = line.split("#")[0].split()
            if len(line) >= 4:
                pts[line[0]] = [float(x) for x in line[1:4]]
    return pts


def band_path(points, nq):
    """k-points of the path (setLoops.F90:38-66): nq steps of equal length
    along the whole path; the k-th point advances along the current
    segment until k * step passes the end of it."""
    P = [np.asarray(p, dtype=float) for p in points]
    seg = [np.linalg.norm(P[i + 1] - P[i]) for i in range(len(P) - 1)]
    step = sum(seg) / nq
    ks, i, j, done = [], 0, 0, 0.0
    d = None
    for kount in range(1, nq + 2):
        if step * kount >= done and i < len(P) - 1:
            i += 1
            done += seg[i - 1]
            j = 0
            d = (P[i] - P[i - 1]) / seg[i - 1]
        ks.append(P[i - 1] + d * j * step)
        j += 1
    return np.array(ks), step


def read_bands(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            if not line.startswith("#") and line.strip():
                rows.append([float(x) for x in line.split()])
    return np.array(rows)


def _case(name):
    deck = D.load_deck(os.path.join(DECKS, name))
    inp = deck["input"]
    bands = inp["band"].split("!")[0].split()
    nq = int(inp["nQvec"])
    pts = read_kbands(os.path.join(DECKS, name, "kbands"))
    m = TBModel(deck["system"])
    assert inp["qbasis"].startswith("b")
    path = [sum(c * m.b[x] for x, c in enumerate(pts[b])) for b in bands]
    ks, step = band_path(path, nq)
    ref = read_bands(glob.glob(os.path.join(DECKS, name, "results_correct", "*", "*Sites", "BS",
                                           "bandstructure_*.dat"))[0])
    return deck, m, ks, step, ref


@pytest.mark.parametrize("name", ["Nb_100_Monolayer", "graphene"])
def test_bands_match_reference(name):
    deck, m, ks, step, ref = _case(name)
    assert len(ks) == len(ref) and ref.shape[1] == m.d + 1
    assert np.allclose(ref[:, 0], step * np.arange(len(ks)), rtol=1e-8, atol=1e-9)
    ev = np.array([np.linalg.eigvalsh(m.hk(k)) for k in ks])
    err = np.abs(ev - ref[:, 1:]).max()
    print(name, "max |band - reference| = %.2e Ry" % err)
    assert err < 5e-8


@pytest.mark.parametrize("name", ["Nb_100_Monolayer", "graphene"])
def test_oracle_bands_match_reference(name):
    """The oracle's own restatement of the set-up (its shells, scaling and
    hee) gives the same bands."""
    deck, m, ks, step, ref = _case(name)
    s = O.OracleSystem(deck["system"])
    ev = np.array([np.linalg.eigvalsh(s.hk(k)) for k in ks[::7]])
    assert np.abs(ev - ref[::7, 1:]).max() < 5e-8


def test_nb_bands_need_the_scaling_law():
    deck, m, ks, step, ref = _case("Nb_100_Monolayer")
    spec = dict(deck["system"])
    spec["types"] = [{k: v for k, v in t.items() if k != "element"} for t in spec["types"]]
    ev = np.array([np.linalg.eigvalsh(TBModel(spec).hk(k)) for k in ks[::20]])
    assert np.abs(ev - ref[::20, 1:]).max() > 1e-3


def test_deck_refuses_what_it_cannot_model(tmp_path):
    src = os.path.join(DECKS, "Nb_100_Monolayer")
    for f in ("basis", "Nb", "kbands"):
        (tmp_path / f).write_text(open(os.path.join(src, f)).read())
    inp = open(os.path.join(src, "input")).read()
    # Un != 0 and no initial occupation file
    (tmp_path / "input").write_text(inp)
    with pytest.raises(ValueError, match="initial occupation"):
        D.load_deck(str(tmp_path))
    for key, val in (("FIELD", "T"), ("superCond", "T"), ("tbmode", "2")):
        (tmp_path / "input").write_text(inp + "\n-> %s = %s\n" % (key, val))
        with pytest.raises(ValueError):
            D.load_deck(str(tmp_path))


def test_fermi_alignment_and_ef_override(tmp_path):
    """Two element files with different Fermi levels: on-site energies move
    to the Fermi level of type 'fermi_layer' (TightBinding.F90:109-115); the
    input key 'Ef' overrides the self-consistency file's E_F."""
    src = os.path.join(DECKS, "nanoribbon")
    for f in ("basis", "A2"):
        (tmp_path / f).write_text(open(os.path.join(src, f)).read())
    a3 = open(os.path.join(src, "A3")).read().splitlines()
    el3 = D.read_element(os.path.join(src, "A3"), 1)
    k = [i for i, l in enumerate(a3) if l.strip() and D._floats(l) == [el3["fermi_level"]]][0]
    a3[k] = " %.6f" % (el3["fermi_level"] + 0.125)
    (tmp_path / "A3").write_text("\n".join(a3) + "\n")
    inp = open(os.path.join(src, "input")).read()
    (tmp_path / "input").write_text(inp.replace("-> fermi_layer = 1", "-> fermi_layer = 2"))
    d = D.load_deck(str(tmp_path))
    el2 = D.read_element(os.path.join(src, "A2"), 1)
    on2, on3 = d["system"]["types"][0]["onsite"], d["system"]["types"][1]["onsite"]
    assert on2 == pytest.approx([v + 0.125 for v in el2["onsite"]])
    assert on3 == pytest.approx(el3["onsite"])
    assert d["system"]["Ef"] == -0.25      # the input's Ef, no self-consistency file here
