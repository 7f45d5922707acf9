"""Self-consistency of the mean-field ground state on the GPU (Newton steps
with the Jacobian from the device's static chi0), and the Goldstone theorem
as its check: without spin-orbit coupling a self-consistent ferromagnet has
a zero mode of the transverse RPA kernel at q = 0, omega = 0, so
(I + chi0 U) becomes singular there (condition number limited only by the
residual of the solution), while the seeded, non-self-consistent state is
better conditioned by orders of magnitude."""
import numpy as np
import pytest

from titan_amd import configs as CF
from titan_amd import selfconsistency as SC
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu


def test_selfconsistent_state_has_a_goldstone_mode():
    """The cfg0 element at U_d = 3 eV (a ferromagnet: about 1.9 muB at this
    mesh), seeded at 2.2 muB, spin-orbit coupling off."""
    spec = CF.bcc_1site(seed=0, U_ev=3.0)
    spec["socscale"] = 0.0
    run = {"nk": (6, 6, 6), "parts": 1, "n1gl": 24, "parts3": 1, "n3gl": 8, "eta": 5e-4, "etap": 5e-4}
    print("seeded state", flush=True)
    _, cond0 = SC.transverse_scan(spec, run, [0.0])
    sc, hist = SC.solve(spec, run, 8.0, tol=1e-11, log=lambda s: print(s, flush=True))
    assert hist[-1] < 1e-11 and len(hist) <= 20
    mz = np.asarray(sc["mz"]).sum()
    print("self-consistent moment %.4f muB, E_F %.6f Ry" % (mz, sc["Ef"]), flush=True)
    assert mz > 1.0
    _, cond = SC.transverse_scan(sc, run, [0.0])
    print("cond at omega = 0: seeded %.3e, self-consistent %.3e" % (cond0[0], cond[0]), flush=True)
    assert cond[0] > 1e8 and cond[0] > 1e3 * cond0[0]


def test_stored_states_are_selfconsistent():
    """The states configs.system applies (sc_states.json) satisfy the
    equations on their configurations' own meshes and contours."""
    for cfg in (2,):
        st = CF.sc_state(cfg)
        if st is None:
            pytest.skip("no stored self-consistent state for cfg%d" % cfg)
        spec, run = CF.system(cfg), CF.run_params(cfg)
        m = TBModel(spec)
        names = SC.unknowns(spec, m)
        K, Y, W, kp, wk = SC.imag_list(m, run)
        h = m.handle()
        h.set_imag_points(K, Y, W)
        ev = h.expectation_values(m.Ef, run["eta"])
        h.close()
        F = SC.residual(spec, m, names, SC.get_x(spec, m, names), ev, CF.filling(cfg))
        print("cfg%d residual %.3e" % (cfg, np.abs(F).max()), flush=True)
        assert np.abs(F).max() < 1e-8
