"""GPU parity of th_chi_traces (orbital traces, spin-orbit cross responses,
rotation to the local frames; calculate_chi.F90:91-185) against the oracle
(oracle/soc.py) on susceptibilities computed by the library itself."""
import numpy as np
import pytest

from oracle import soc
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd.system import TBModel, l_matrices

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _lvec(model):
    L = np.transpose(l_matrices(), (1, 2, 0))       # [mu, gamma, s] over orbital ids 1..9
    o = model.orbs - 1
    return L[np.ix_(o, o, [0, 1, 2])]


@pytest.mark.parametrize("which", ["cfg0", "bcc2", "ribbon"])
def test_traces_match_oracle(which):
    spec = {"cfg0": lambda: CF.system(0), "bcc2": lambda: CF.bcc_supercell(2),
            "ribbon": lambda: CF.system(1)}[which]()
    m = TBModel(spec)
    nk = 2
    y, wy = Ct.imag_epoints(1, 4)
    dims = (nk, nk, 1) if which == "ribbon" else (nk, nk, nk)
    kp, wk = Ct.uniform_kmesh(m.b, *dims)
    h = m.handle()
    h.set_imag_points(np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
    h.set_real_kmesh(kp, wk)
    om = 4e-3
    x2, p2 = Ct.real_epoints(m.Ef, om, 1, 4)
    chi0 = h.chi0([0.02, -0.01, 0.0], om, m.Ef, 1e-3, 1e-3, x2, p2)
    chi, _ = h.dyson(chi0, m.Umat)
    lv = _lvec(m)
    rng = np.random.default_rng(7)
    ang = np.column_stack([rng.uniform(0, 180, m.nA), rng.uniform(-180, 180, m.nA)])
    for c in (chi0, chi):
        s_ref = soc.spin_traces(m.nA, m.nOrb, c)
        ls_ref, sl_ref, ll_ref = soc.soc_traces(m.nA, m.nOrb, c, lv)
        r_ref = soc.rotate_spin_traces(m.nA, s_ref, ang)
        s, ls, sl, ll = h.traces(c, lv)
        sr, _, _, _ = h.traces(c, angles=ang)
        errs = [rel(s, s_ref), rel(ls, ls_ref), rel(sl, sl_ref), rel(ll, ll_ref), rel(sr, r_ref)]
        print(which, "spin %.1e LS %.1e SL %.1e LL %.1e rotated %.1e" % tuple(errs), flush=True)
        assert max(errs) < 1e-12
        assert np.linalg.norm(ls_ref) > 0 and np.linalg.norm(ll_ref) > 0
    # chi = None: the handle's chi of the last Dyson step, no transfer
    s2, ls2, sl2, ll2 = h.traces(None, lv)
    s1, ls1, sl1, ll1 = h.traces(chi, lv)
    assert rel(s2, s1) < 1e-15 and rel(ls2, ls1) < 1e-15 and rel(ll2, ll1) < 1e-15
    h.close()
