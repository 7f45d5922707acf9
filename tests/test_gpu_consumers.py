"""The reference's host consumers of chi fed with the device's chi: the
eigen-decomposition of chi^{+-} (diagonalize_susceptibilities,
mod_susceptibilities.F90:162-189), the Gilbert damping rows (write_alpha,
mod_alpha.F90:128-221) and J, K from the static susceptibility
(get_J_K_from_chi, mod_coupling.F90:130-241), on a two-atom cell.  The same
consumers fed with the FP64 oracle's chi (eintshechi, the Dyson step and the
orbital trace restated in C) and moments (expectation_values_greenfunction
restated in numpy) are the reference: every consumer output within 1e-8 of
it (the GPU's chi is FP64 within 1e-10 of the oracle's)."""
import numpy as np
import pytest

from oracle import expectation as OE
from oracle import oracle as O
from titan_amd import alpha as A
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd import coupling as Cp
from titan_amd import titan_io as IO
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu


def _close(a, b, tol=1e-8):
    a, b = np.asarray(a), np.asarray(b)
    return np.abs(a - b).max() <= tol * max(1.0, np.abs(b).max())


def test_host_consumers_on_device_chi():
    spec = CF.bcc_supercell(2, seed=3)
    m, s = TBModel(spec), O.OracleSystem(spec)
    nA, nO = m.nA, m.nOrb
    y, wy = Ct.imag_epoints(1, 8)
    kp, wk = Ct.uniform_kmesh(m.b, 3, 3, 3)
    K, Y, W = np.tile(kp, (len(y), 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel()
    eta = 5e-4
    oms = [0.0, 2e-3, 4e-3]
    h = m.handle()
    h.set_imag_points(K, Y, W)
    h.set_real_kmesh(kp, wk)
    print("device sweep, nA", nA, flush=True)
    s0, sc, _ = h.sweep([0.0, 0.0, 0.0], oms, m.Ef, eta, eta, 1, 6, m.Umat)
    ev = h.expectation_values(m.Ef, eta)
    h.close()
    U = s.umat()
    o0, oc = [], []
    for om in oms:
        x2, p2 = Ct.real_epoints(m.Ef, om, 1, 6)
        c0 = s.chi0([0.0, 0.0, 0.0], om, eta, eta, (K, Y, W), (kp, wk), (x2, p2))
        o0.append(O.trace(nA, nO, c0))
        oc.append(O.trace(nA, nO, O.dyson(c0, U)))
    for w in range(len(oms)):
        assert _close(sc[w], oc[w], 1e-10) and _close(s0[w], o0[w], 1e-10)
    ov = OE.expectation_values(s, K, Y, W, m.Ef, eta)
    # eigen-decomposition of chi^{+-} at every frequency
    for w in range(len(oms)):
        eg, _ = IO.diagonalize_susceptibilities(sc[w], nA)
        eo, _ = IO.diagonalize_susceptibilities(oc[w], nA)
        assert _close(np.sort_complex(eg), np.sort_complex(eo))
    # damping rows at omega != 0
    mvec = np.stack([ev[1].sum(axis=1), ev[2].sum(axis=1), ev[3].sum(axis=1)], axis=1)
    mvo = np.stack([ov[1].sum(axis=1), ov[2].sum(axis=1), ov[3].sum(axis=1)], axis=1)
    assert _close(mvec, mvo, 1e-10)
    mabs = np.sqrt((mvec ** 2).sum(axis=1))
    Um = [spec["types"][t]["U"][1] for t in spec["type"]]
    for w in (1, 2):
        rg, vg = A.alpha_rows(oms[w], sc[w], s0[w], mabs, Um)
        ro, vo = A.alpha_rows(oms[w], oc[w], o0[w], np.sqrt((mvo ** 2).sum(axis=1)), Um)
        for k in vg:
            assert _close(vg[k], vo[k]), k
        assert sorted(rg) == sorted(ro)
    # J and K from the static susceptibility
    jg = Cp.get_J_K_from_chi(sc[0], mvec)
    jo = Cp.get_J_K_from_chi(oc[0], mvo)
    for k in ("J", "Jtotal", "Kx", "Ky"):
        assert _close(jg[k], jo[k]), k
    print("J(1,2) = %.6e meV, Kx(1) = %.6e meV" % (jg["J"][0, 1], jg["Kx"][0]), flush=True)
