"""GPU checks of the adjoint (th_chi_adjoint) and of the parameter gradients
of titan_amd.autodiff.

  * th_chi_adjoint in FP64 against the FP64 adjoint reference
    (oracle/adjoint.py, itself pinned to finite differences of the C oracle
    by tests/test_adjoint_oracle.py): t_hat, h0_hat and U_hat to 1e-10
    relative Frobenius; in mixed precision to 1e-5.
  * gradients with respect to U_d, lambda_d, Slater-Koster integrals and an
    on-site energy through torch.autograd against central finite differences
    (h = 1e-5, five-point stencil) of the C oracle's chi built from the
    perturbed parameter set, to 1e-6 relative.
  * bitwise reproducibility of the adjoint; the three-product contraction;
    the Green functions by inversion (Hhat from each pair's G instead of the
    spectral form).
  Every Green function is recomputed in the backward by default; a leg of
  each precision reads the imaginary-axis B side from the forward's panel
  cache instead (th_tune "adjoint_bcache"), after a forward of that
  precision.
"""
import copy

import numpy as np
import pytest
import torch

from oracle import adjoint as AJ
from oracle import oracle as O
from titan_amd import binding as B
from titan_amd import configs as CF
from titan_amd import contour as Ct
from titan_amd.system import TBModel

pytestmark = pytest.mark.gpu


def _setup(spec, nk, n1, n3, omega, q=(0.0, 0.0, 0.0), eta=5e-3):
    m = TBModel(spec)
    h = m.handle()
    y, wy = Ct.imag_epoints(1, n1)
    kp, wk = Ct.uniform_kmesh(m.b, *(nk if a < m.sysdim else 1 for a in range(3)))
    imag = (np.tile(kp, (n1, 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
    h.set_imag_points(*imag)
    h.set_real_kmesh(kp, wk)
    x2, p2 = Ct.real_epoints(m.Ef, omega, 1, n3) if abs(omega) >= 1e-15 else (None, None)
    return dict(m=m, h=h, imag=imag, real_k=(kp, wk), x2=x2, p2=p2, omega=omega, q=np.asarray(q, float), eta=eta)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


CASES = {
    # name: (spec, nk, n1gl, n3gl, omega, q, with U)
    "bcc1_real_q_U": (CF.bcc_1site, 3, 6, 6, 3e-3, (0.05, -0.02, 0.1), True),
    "bcc1_imag_U": (CF.bcc_1site, 3, 6, 6, 0.0, (0.0, 0.0, 0.0), True),
    # 20 real nodes: the real pairs' grouped order (blocks of 16 nodes, then 4)
    "bcc1_real_nx20_U": (CF.bcc_1site, 2, 4, 20, 3e-3, (0.05, -0.02, 0.1), True),
    # 40 energies on 2^3 k-points: three runs per k-point (16, 16, 8) in one batch,
    # so the per-k-point sums of the spectral form take three launch rounds
    "bcc1_imag_n40_U": (CF.bcc_1site, 2, 40, 4, 0.0, (0.0, 0.0, 0.0), True),
    "bcc2_real_chi0": (lambda: CF.bcc_supercell(2), 2, 4, 4, 2e-3, (0.0, 0.0, 0.07), False),
    "bcc2_imag_U": (lambda: CF.bcc_supercell(2), 2, 4, 4, 0.0, (0.0, 0.0, 0.0), True),
    "pchain_real_U0": (lambda: CF.chain_p(), 5, 6, 6, 2e-3, (0.1, 0.0, 0.0), True),
    "bcc3_real_U": (lambda: CF.system(2), 1, 3, 3, 3e-3, (0.0, 0.0, 0.0), True),
    "bcc4_imag_chi0": (lambda: CF.bcc_supercell(4), 1, 3, 3, 0.0, (0.0, 0.0, 0.0), False),
}


@pytest.mark.parametrize("name", list(CASES))
def test_adjoint_matches_fp64_reference(name):
    mk, nk, n1, n3, omega, q, useU = CASES[name]
    spec = mk()
    print("case", name, flush=True)
    c = _setup(spec, nk, n1, n3, omega, q)
    m, h = c["m"], c["h"]
    chi0 = h.chi0(c["q"], omega, m.Ef, c["eta"], c["eta"], c["x2"], c["p2"])
    U = m.Umat if useU else None
    chi = h.dyson(chi0, U)[0] if useU else None
    rng = np.random.default_rng(5)
    W = rng.standard_normal((m.D, m.D)) + 1j * rng.standard_normal((m.D, m.D))
    th, hh, uh = h.adjoint(c["q"], omega, m.Ef, c["eta"], c["eta"], chi0, W, U=U, chi=chi, x2=c["x2"], p2=c["p2"])
    am = AJ.ArrayModel(m.nA, m.nOrb, m.R, m.jb, m.ib, m.t, m.h0)
    re_k = c["real_k"] if c["x2"] is not None else None
    re_e = (c["x2"], c["p2"]) if c["x2"] is not None else None
    rt, rh, ru = AJ.adjoint(am, c["q"], omega, m.Ef, c["eta"], c["eta"], c["imag"], re_k, re_e, chi0, W, U=U, chi=chi)
    print("rel t %.2e h0 %.2e" % (_rel(th, rt), _rel(hh, rh)), flush=True)
    assert _rel(th, rt) <= 1e-10
    assert _rel(hh, rh) <= 1e-10
    if useU:
        assert _rel(uh, ru) <= 1e-10
    # the imaginary-axis B side from the FP64 forward's panel cache
    h.tune(adjoint_bcache=1)
    thc, hhc, _ = h.adjoint(c["q"], omega, m.Ef, c["eta"], c["eta"], chi0, W, U=U, chi=chi, x2=c["x2"], p2=c["p2"])
    assert _rel(thc, rt) <= 1e-10 and _rel(hhc, rh) <= 1e-10
    # ... and from a mixed forward's (split half-precision words)
    h.chi0(c["q"], omega, m.Ef, c["eta"], c["eta"], c["x2"], c["p2"], precision=B.PREC_MIXED)
    thm, hhm, _ = h.adjoint(c["q"], omega, m.Ef, c["eta"], c["eta"], chi0, W, U=U, chi=chi, x2=c["x2"],
                            p2=c["p2"], precision=B.PREC_MIXED)
    assert _rel(thm, rt) <= 1e-5 and _rel(hhm, rh) <= 1e-5
    h.tune(adjoint_bcache=0)
    # mixed precision: FP32 panels in the two adjoint contractions
    th2, hh2, uh2 = h.adjoint(c["q"], omega, m.Ef, c["eta"], c["eta"], chi0, W, U=U, chi=chi, x2=c["x2"],
                              p2=c["p2"], precision=B.PREC_MIXED)
    assert _rel(th2, rt) <= 1e-5
    assert _rel(hh2, rh) <= 1e-5
    # bitwise reproducible
    th3, hh3, _ = h.adjoint(c["q"], omega, m.Ef, c["eta"], c["eta"], chi0, W, U=U, chi=chi, x2=c["x2"], p2=c["p2"])
    assert np.array_equal(th3, th) and np.array_equal(hh3, hh)
    # three real MFMA products per complex one in the adjoint contractions
    h.tune(contract_3m=1)
    th4, hh4, _ = h.adjoint(c["q"], omega, m.Ef, c["eta"], c["eta"], chi0, W, U=U, chi=chi, x2=c["x2"], p2=c["p2"])
    assert _rel(th4, rt) <= 1e-10 and _rel(hh4, rh) <= 1e-10
    # Green functions by inversion: Hhat from each pair's G (the spectral
    # form V [sum_p ...] V^H needs the eigenpairs), both precisions
    h.tune(contract_3m=0, spectral=0)
    th5, hh5, _ = h.adjoint(c["q"], omega, m.Ef, c["eta"], c["eta"], chi0, W, U=U, chi=chi, x2=c["x2"], p2=c["p2"])
    assert _rel(th5, rt) <= 1e-10 and _rel(hh5, rh) <= 1e-10
    th6, hh6, _ = h.adjoint(c["q"], omega, m.Ef, c["eta"], c["eta"], chi0, W, U=U, chi=chi, x2=c["x2"],
                            p2=c["p2"], precision=B.PREC_MIXED)
    assert _rel(th6, rt) <= 1e-5 and _rel(hh6, rh) <= 1e-5
    h.close()


def _set_param(spec, key, delta):
    ty = spec["types"][0]
    if key[0] == "U":
        ty["U"][key[1]] += delta
    elif key[0] == "lam":
        ty["lambda"][key[1]] += delta
    elif key[0] == "sk":
        ty["sk"][key[1]][key[2]] += delta
    else:
        ty["onsite"][key[1]] += delta


def test_parameter_gradients_match_finite_differences():
    from titan_amd.autodiff import ChiOperator, TorchModel
    spec = CF.bcc_1site()
    nk, n1, n3, eta = 3, 8, 8, 5e-3
    oms = [0.0, 4e-3]
    rng = np.random.default_rng(3)
    Wt = [rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)) for _ in oms]
    tm = TorchModel(spec)
    op = ChiOperator(tm, (nk, nk, nk), n1gl=n1, n3gl=n3, eta=eta)
    p = tm.params()
    L = 0.0
    for W, om in zip(Wt, oms):
        s = op.trace(op(p, om))
        L = L + torch.real(torch.sum(torch.from_numpy(W).conj() * s))
    L.backward()
    kp, wk = op.kp, op.wk
    imag = (op.kim, op.yim, op.wim)

    def L_oracle(sp):
        so = O.OracleSystem(sp)
        U = so.umat()
        tot = 0.0
        for W, om in zip(Wt, oms):
            re_k = (kp, wk) if om else None
            re_e = Ct.real_epoints(so.Ef, om, 1, n3) if om else None
            c0 = so.chi0([0, 0, 0], om, eta, eta, imag, re_k, re_e)
            tot += np.real(np.vdot(W, O.trace(1, tm.m.nOrb, O.dyson(c0, U))))
        return tot

    Lv = float(L.detach())
    assert abs(Lv - L_oracle(spec)) <= 1e-10 * abs(Lv)
    h = 1e-5
    checks = ((("U", 1), p["U"].grad[0, 1]), (("lam", 1), p["lam"].grad[0, 1]),
              (("sk", 0, 4), p["sk"].grad[0, 0, 4]), (("sk", 0, 0), p["sk"].grad[0, 0, 0]),
              (("sk", 1, 8), p["sk"].grad[0, 1, 8]), (("on", 4), p["onsite"].grad[0, 4]))
    for key, g in checks:
        vals = []
        for k in (2, 1, -1, -2):
            sp = copy.deepcopy(spec)
            _set_param(sp, key, k * h)
            vals.append(L_oracle(sp))
        fd = (-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * h)
        print(key, float(g), fd, flush=True)
        assert abs(float(g) - fd) <= 1e-6 * abs(fd), (key, float(g), fd)


def test_cfg3_loss_gradients():
    """The adjoint workload's loss, sum over its 32 frequencies of
    (Im chi^{+-}(omega) - seeded target)^2 on the 1-site bcc system (reduced
    mesh: 3^3 k-points, 8 + 8 energies): the value against the oracle; the
    gradients with respect to U_d, lambda_d and two Slater-Koster integrals
    against five-point central differences (h = 1e-5) of the oracle's loss
    (1e-6 relative); the mixed-precision gradient of U_d, lambda_d and the
    ten first-shell integrals within 1e-5 of the FP64 one."""
    from titan_amd.autodiff import ChiOperator, TorchModel, chi_pm_loss, nominal_chi_pm, target_curve
    spec = CF.system(3)
    run = CF.run_params(3)
    oms = list(run["omega"])
    assert len(oms) == 32
    nk, n1, n3, eta = 3, 8, 8, 5e-3
    tm = TorchModel(spec)
    nom = nominal_chi_pm(ChiOperator(tm, (nk, nk, nk), n1gl=n1, n3gl=n3, eta=eta), tm, oms)
    tgt = target_curve(oms, np.max(np.abs(nom)))
    print("largest |Im chi+-| %.4e" % np.max(np.abs(nom)), flush=True)
    grads = {}
    for prec in (B.PREC_FP64, B.PREC_MIXED):
        op = ChiOperator(tm, (nk, nk, nk), n1gl=n1, n3gl=n3, eta=eta, precision=prec)
        p = tm.params()
        L = chi_pm_loss(op, p, oms, tgt)
        L.backward()
        grads[prec] = (float(L.detach()), p)
    Lv, p = grads[B.PREC_FP64]
    op = ChiOperator(tm, (nk, nk, nk), n1gl=n1, n3gl=n3, eta=eta)
    kp, wk = op.kp, op.wk
    imag = (op.kim, op.yim, op.wim)

    def L_oracle(sp):
        so = O.OracleSystem(sp)
        U = so.umat()
        tot = 0.0
        for om, tg in zip(oms, tgt):
            re_k = (kp, wk) if om else None
            re_e = Ct.real_epoints(so.Ef, om, 1, n3) if om else None
            c0 = so.chi0([0, 0, 0], om, eta, eta, imag, re_k, re_e)
            s = O.trace(1, tm.m.nOrb, O.dyson(c0, U))
            tot += (s[0, 0].imag - tg) ** 2
        return tot

    Lo = L_oracle(spec)
    print("loss GPU %.12e oracle %.12e" % (Lv, Lo), flush=True)
    assert abs(Lv - Lo) <= 1e-10 * abs(Lo)
    h = 1e-5
    checks = ((("U", 1), p["U"].grad[0, 1]), (("lam", 1), p["lam"].grad[0, 1]),
              (("sk", 0, 4), p["sk"].grad[0, 0, 4]), (("sk", 0, 9), p["sk"].grad[0, 0, 9]))
    for key, g in checks:
        vals = []
        for k in (2, 1, -1, -2):
            sp = copy.deepcopy(spec)
            _set_param(sp, key, k * h)
            vals.append(L_oracle(sp))
        fd = (-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * h)
        print(key, float(g), fd, flush=True)
        assert abs(float(g) - fd) <= 1e-6 * abs(fd), (key, float(g), fd)
    pm = grads[B.PREC_MIXED][1]
    g64 = np.concatenate([[float(p["U"].grad[0, 1]), float(p["lam"].grad[0, 1])], p["sk"].grad[0, 0].numpy()])
    gmx = np.concatenate([[float(pm["U"].grad[0, 1]), float(pm["lam"].grad[0, 1])], pm["sk"].grad[0, 0].numpy()])
    print("mixed vs FP64 gradient %.2e" % _rel(gmx, g64), flush=True)
    assert _rel(gmx, g64) < 1e-5
