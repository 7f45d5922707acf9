"""CPU checks of the FP64 adjoint reference (oracle/adjoint.py).

The adjoint has no counterpart in the reference, so it is pinned by the
mathematics: its directional derivative Re<xhat, dx> must equal a central
finite difference (h = 1e-5) of the C oracle's forward chi0 / Dyson (which
the fixture tests pin to the reference's own outputs), in random unit
directions of the hop table t, of h0 and of U, to 1e-8 of the gradient norm
(the Cauchy-Schwarz bound of the directional derivative; the error of the
central difference at this h is far below it).  A planted mistake (the
conjugate term of the imaginary-axis contraction put on the wrong entries)
must fail the same check.
"""
import numpy as np
import pytest

from oracle import adjoint as AJ
from oracle import oracle as O
from titan_amd import configs as CF


def _case(spec, nk=2, n1=4, n3=4, omega=3e-3, q=(0.0, 0.0, 0.0)):
    s = O.OracleSystem(spec)
    R, jb, ib, t, h0, U, b = s.tables()
    kp, wk = s.kmesh_uniform(nk, nk, nk)
    y, wy = O.imag_epoints(1, n1)
    imag = (np.tile(kp, (n1, 1)), np.repeat(y, len(wk)), np.outer(wy, wk).ravel())
    x2, p2 = O.real_epoints(s.Ef, omega, 1, n3)
    return dict(s=s, R=R, jb=jb, ib=ib, t=t, h0=h0, U=U, b=b, imag=imag, real_k=(kp, wk), real_e=(x2, p2),
                omega=omega, q=np.asarray(q, dtype=float), eta=5e-3, etap=5e-3)


def _forward_c(c, t, h0, U):
    """chi0 and chi from the C oracle for explicit tables."""
    s = O.OracleSystem.from_arrays(c["s"].nA, c["s"].nOrb, c["R"], c["jb"], c["ib"], t, h0, c["b"], c["s"].Ef)
    chi0 = s.chi0(c["q"], c["omega"], c["eta"], c["etap"], c["imag"], c["real_k"], c["real_e"])
    return chi0, (O.dyson(chi0, U) if U is not None else None)


def test_numpy_forward_matches_c_oracle():
    for spec, q in ((CF.bcc_1site(), (0.1, 0.0, 0.05)), (CF.bcc_supercell(2), (0.0, 0.07, 0.0))):
        c = _case(spec, q=q)
        m = AJ.ArrayModel(c["s"].nA, c["s"].nOrb, c["R"], c["jb"], c["ib"], c["t"], c["h0"])
        ref, _ = _forward_c(c, c["t"], c["h0"], None)
        mine = AJ.chi0(m, c["q"], c["omega"], c["s"].Ef, c["eta"], c["etap"], c["imag"], c["real_k"], c["real_e"])
        assert np.linalg.norm(mine - ref) <= 1e-13 * np.linalg.norm(ref)


def _directional(c, use_U, seed, drop_conj=False):
    rng = np.random.default_rng(seed)
    s = c["s"]
    D, d = s.D, s.d
    unit = lambda a: a / np.linalg.norm(a)
    W = rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))     # L = Re sum conj(W) x
    dt = unit(rng.standard_normal(c["t"].shape))                           # hops stay real
    dh = unit(rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d)))
    dU = unit((rng.standard_normal((D, D)) + 1j * rng.standard_normal((D, D))) * (c["U"] != 0)) if use_U else None
    U = c["U"] if use_U else None
    chi0, chi = _forward_c(c, c["t"], c["h0"], U)
    m = AJ.ArrayModel(s.nA, s.nOrb, c["R"], c["jb"], c["ib"], c["t"], c["h0"])
    if drop_conj:
        th, hh, uh = _adjoint_without_conj_term(m, c, chi0, W, U, chi)
    else:
        th, hh, uh = AJ.adjoint(m, c["q"], c["omega"], s.Ef, c["eta"], c["etap"], c["imag"], c["real_k"],
                                c["real_e"], chi0, W, U=U, chi=chi)
    pred = np.sum(th.real * dt) + np.real(np.vdot(hh, dh))
    scale = np.linalg.norm(th.real) + np.linalg.norm(hh)     # |pred| <= scale for unit directions
    if use_U:
        pred += np.real(np.vdot(uh, dU))
        scale += np.linalg.norm(uh * (c["U"] != 0))
    h = 1e-5

    def loss(sign):
        x0, x = _forward_c(c, c["t"] + sign * h * dt, c["h0"] + sign * h * dh,
                           (U + sign * h * dU) if use_U else None)
        return np.real(np.vdot(W, x if use_U else x0))
    fd = (loss(1) - loss(-1)) / (2 * h)
    return pred, fd, scale


def _adjoint_without_conj_term(m, c, chi0, W, U, chi):
    """The adjoint with the conj(G(i,j) Gq(j,i)) term of the imaginary axis
    left out of the chi0 -> accumulator step (a planted mistake)."""
    orig = AJ.chi_index_maps

    def maps(nA, nOrb):
        (b1, r1, c1), (b2, r2, c2) = orig(nA, nOrb)
        return (b1, r1, c1), (b1, r1, c1)     # term 2 lands on term 1's entries instead
    AJ.chi_index_maps = maps
    try:
        return AJ.adjoint(m, c["q"], c["omega"], c["s"].Ef, c["eta"], c["etap"], c["imag"], c["real_k"],
                          c["real_e"], chi0, W, U=U, chi=chi)
    finally:
        AJ.chi_index_maps = orig


@pytest.mark.parametrize("which", ["chi0_imag", "chi_real_q", "chi0_two_atoms"])
def test_adjoint_reference_matches_finite_differences(which):
    if which == "chi0_imag":
        c, use_U = _case(CF.bcc_1site(), omega=0.0), False
    elif which == "chi_real_q":
        c, use_U = _case(CF.bcc_1site(), q=(0.05, -0.02, 0.1)), True
    else:
        c, use_U = _case(CF.bcc_supercell(2), nk=2, n1=3, n3=3, q=(0.0, 0.0, 0.08)), False
    for seed in (11, 12):
        pred, fd, scale = _directional(c, use_U, seed=seed)
        assert abs(pred - fd) <= 1e-8 * scale, (pred, fd, scale)


def test_planted_mistake_fails_the_check():
    c = _case(CF.bcc_1site(), omega=0.0)
    pred, fd, scale = _directional(c, False, seed=11, drop_conj=True)
    assert abs(pred - fd) > 1e-3 * scale
